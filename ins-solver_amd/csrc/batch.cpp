#include "batch.h"

#include "amg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <stdexcept>

#include "dg_layout.h"
#include "kernels.h"
#include "refelem.h"
#include "solver.h"
#include "vec_kernels.h"

namespace insp {

namespace {
int64_t round_up(int64_t v, int64_t m) { return ((v + m - 1) / m) * m; }
size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
int np_of(int dim, int N) { return dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6; }
int nfp_of(int dim, int N) { return dim == 2 ? N + 1 : (N + 1) * (N + 2) / 2; }
}  // namespace

// Autotuned tile size per (dimension, V-cycle precision), shared by the engines of a process.
std::mutex g_tune_mu;
std::map<std::pair<int, bool>, BatchEngine::TuneResult> g_tune;

struct BatchEngine::Prob {
  int slot = -1, N = 0, Np = 0, Np1 = 0, Nfp = 0, dim = 0;
  int64_t E = 0, S = 0, nbf = 0;
  bool var = false, nullspace = false;
  double rtol = 1e-8, atol = 0, tol = 0, bnorm = 0;
  int max_iter = 500, iters = 0;
  void* arena = nullptr;
  size_t bytes = 0;
  BatchProb desc{};
  ProbResult result;
  bool done = false;
  std::chrono::steady_clock::time_point t0;
  void* amg_arena = nullptr;     // vertex maps + hierarchy (AMG coarse solve)
  size_t amg_bytes = 0;
  std::vector<int> amg_rows;     // rows per AMG level
};

struct BatchEngine::View {
  // the view's index tables in one device buffer, filled from one pinned host image per refresh:
  // tiles (int2, tile_cap), probs (int, prob_cap), ranges (int2, max slots), counts (int, 12),
  // AMG row tiles (int2, levels x ltile_cap), AMG tile counts (int, levels)
  DBuf<char> meta;
  char* h_meta = nullptr;
  size_t meta_bytes = 0;
  size_t o_tiles = 0, o_probs = 0, o_ranges = 0, o_counts = 0, o_ltiles = 0, o_lcount = 0;
  DBuf<double> partial;
  int tile_cap = 0, prob_cap = 0, ltile_cap = 0;
  int levels = 0;                // deepest AMG hierarchy in the view
  int ocap[5] = {0, 0, 0, 0, 0};  // grid per order
  BatchView bv{};
  template <typename X>
  X* host(size_t off) { return reinterpret_cast<X*>(h_meta + off); }
  template <typename X>
  X* dev(size_t off) { return reinterpret_cast<X*>(meta.p + off); }
  ~View() {
    if (h_meta) (void)hipHostFree(h_meta);
  }
};

size_t BatchEngine::arena_bytes(int dim, int64_t cells, int N, bool var, bool fp32) {
  const int64_t S = round_up(std::max<int64_t>(cells, 1), kTile);
  const int Np = np_of(dim, N), Np1 = dim + 1, Nf = dim + 1, Nfp = nfp_of(dim, N);
  const int NK = dim == 2 ? 3 : 6;
  const size_t d = 8, t = fp32 ? 4 : 8;
  size_t b = 0;
  b += align256(NK * S * d) + align256(S * d);                    // Kt, J
  b += 2 * align256(Nf * S * 4);                                  // nbr, code
  b += align256((size_t)std::max<int64_t>(cells, 1) * Nf * 4);    // bctype (upper bound)
  b += 2 * align256(Nf * S * d);                                  // tauF, tauC
  if (var) b += align256(Np * S * d) + align256(Np * S * 4) + align256(Np1 * S * d) + align256(Np1 * S * 4);
  b += align256(Np * S * d) + align256(Np * S * 4) + align256(Np1 * S * d) + align256(Np1 * S * 4);  // inverses
  b += 6 * align256(Np * S * d) + align256((size_t)Nf * Nfp * S * d);  // outer FCG + traces
  b += 4 * align256(Np * S * t) + (fp32 ? align256((size_t)Nf * Nfp * S * t) : 0);
  b += 6 * align256(Np1 * S * t);
  return b;
}

BatchEngine::BatchEngine(const BatchConfig& cfg, hipStream_t stream) : cfg_(cfg), stream_(stream) {
  if (cfg_.dim != 2 && cfg_.dim != 3) throw InvalidArgument("batch dimension must be 2 or 3");
  if (cfg_.cheb_order < 1 || cfg_.cheb_order > 8 || cfg_.pre_it < 1 || cfg_.post_it < 0)
    throw InvalidArgument("invalid batch smoother settings");
  if (cfg_.tile != 0 && cfg_.tile != 64 && cfg_.tile != 128 && cfg_.tile != 256)
    throw InvalidArgument("tile must be 64, 128, 256 or 0 (autotuned)");
  if (cfg_.tile) {
    tile_ = cfg_.tile;
    tuned_ = true;
  } else {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find({cfg_.dim, cfg_.fp32});
    if (it != g_tune.end()) {
      tile_ = it->second.tile;
      tuned_ = true;
    }
  }
  if (!stream_) {
    INSP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  max_slots_ = 4096;
  INSP_HIP(hipMalloc(&d_desc_, sizeof(BatchProb) * max_slots_));
  INSP_HIP(hipMalloc(&d_scal_, sizeof(double) * BS_COUNT * max_slots_));
  INSP_HIP(hipMemset(d_scal_, 0, sizeof(double) * BS_COUNT * max_slots_));
  INSP_HIP(hipDeviceSynchronize());
  INSP_HIP(hipHostMalloc((void**)&h_scal_, sizeof(double) * BS_COUNT * max_slots_, hipHostMallocDefault));
  INSP_HIP(hipHostMalloc((void**)&h_stage_, sizeof(double) * 5 * max_slots_, hipHostMallocDefault));
  INSP_HIP(hipMalloc(&d_stage_, sizeof(double) * 5 * max_slots_));
  probs_.resize(max_slots_);
  for (int i = max_slots_ - 1; i >= 0; --i) free_slots_.push_back(i);
  run_ = std::make_unique<View>();
  work_ = std::make_unique<View>();
  build_tables();
}

BatchEngine::~BatchEngine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  if (h_stage_) (void)hipHostFree(h_stage_);
  if (d_stage_) (void)hipFree(d_stage_);
  for (auto& p : probs_) {
    if (p && p->arena) (void)hipFree(p->arena);
    if (p && p->amg_arena) (void)hipFree(p->amg_arena);
  }
  for (void* b : table_bufs_) (void)hipFree(b);
  if (d_desc_) (void)hipFree(d_desc_);
  if (d_scal_) (void)hipFree(d_scal_);
  if (h_scal_) (void)hipHostFree(h_scal_);
  run_.reset();
  work_.reset();
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

void BatchEngine::build_tables() {
  const int d = cfg_.dim;
  auto up = [&](const void* h, size_t bytes) {
    void* p = nullptr;
    INSP_HIP(hipMalloc(&p, bytes));
    INSP_HIP(hipMemcpy(p, h, bytes, hipMemcpyHostToDevice));
    table_bufs_.push_back(p);
    return p;
  };
  const RefElement r1 = make_ref_element(d, 1);
  for (int N = 1; N <= kMaxOrder; ++N) {
    const RefElement re = make_ref_element(d, N);
    std::vector<double> C;
    std::vector<int> tab;
    make_op_tables(re, C, tab);
    std::vector<float> Cf(C.begin(), C.end());
    tables_.C64[N] = (const double*)up(C.data(), C.size() * 8);
    tables_.C32[N] = (const float*)up(Cf.data(), Cf.size() * 4);
    tables_.tab[N] = (const int*)up(tab.data(), tab.size() * 4);
    if (N > 1) {
      const std::vector<double> P = interp_matrix(re, r1);
      const std::vector<double> Pt = mat_transpose(P, re.Np, r1.Np);
      std::vector<float> Pf(P.begin(), P.end()), Ptf(Pt.begin(), Pt.end());
      tables_.P64[N] = (const double*)up(P.data(), P.size() * 8);
      tables_.Pt64[N] = (const double*)up(Pt.data(), Pt.size() * 8);
      tables_.P32[N] = (const float*)up(Pf.data(), Pf.size() * 4);
      tables_.Pt32[N] = (const float*)up(Ptf.data(), Ptf.size() * 4);
    }
  }
}

void BatchEngine::upload_desc(int slot) {
  INSP_HIP(hipMemcpyAsync(d_desc_ + slot, &probs_[slot]->desc, sizeof(BatchProb),
                          hipMemcpyHostToDevice, stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::set_view(View& v, const std::vector<int>& slots) {
  int ntiles = 0;
  for (int s : slots) ntiles += (int)(probs_[s]->S / tile_);
  const int nprobs = (int)slots.size();
  auto grow = [](int need, int cap) {
    int c = std::max(cap, 64);
    while (c < need) c *= 2;
    return c;
  };
  // AMG row tiles per level (256 rows each, independent of the cell tile)
  std::vector<int> per_level(kAmgLevels, 0);
  int levels = 0;
  for (int sl : slots) {
    const Prob& P = *probs_[sl];
    levels = std::max(levels, (int)P.amg_rows.size());
    for (size_t k = 0; k < P.amg_rows.size(); ++k) per_level[k] += (P.amg_rows[k] + kTile - 1) / kTile;
  }
  const int lneed = std::max(1, *std::max_element(per_level.begin(), per_level.end()));
  const int tcap = grow(ntiles, v.tile_cap), pcap = grow(nprobs, v.prob_cap), lcap = grow(lneed, v.ltile_cap);
  if (tcap != v.tile_cap || !v.partial.p) v.partial.alloc((size_t)4 * tcap);
  if (tcap != v.tile_cap || pcap != v.prob_cap || lcap != v.ltile_cap || !v.meta.p) {
    size_t off = 0;
    auto seg = [&](size_t bytes) {
      const size_t at = off;
      off += align256(bytes);
      return at;
    };
    v.o_tiles = seg(sizeof(int2) * tcap);
    v.o_probs = seg(sizeof(int) * pcap);
    v.o_ranges = seg(sizeof(int2) * max_slots_);
    v.o_counts = seg(sizeof(int) * 12);
    v.o_ltiles = seg(sizeof(int2) * kAmgLevels * lcap);
    v.o_lcount = seg(sizeof(int) * kAmgLevels);
    v.meta_bytes = off;
    v.meta.alloc(off);
    if (v.h_meta) (void)hipHostFree(v.h_meta);
    v.h_meta = nullptr;
    INSP_HIP(hipHostMalloc((void**)&v.h_meta, off, hipHostMallocDefault));
    std::memset(v.h_meta, 0, off);
    v.tile_cap = tcap;
    v.prob_cap = pcap;
    v.ltile_cap = lcap;
  }
  int2* h_tiles = v.host<int2>(v.o_tiles);
  int* h_probs = v.host<int>(v.o_probs);
  int2* h_ranges = v.host<int2>(v.o_ranges);
  int* h_counts = v.host<int>(v.o_counts);
  int2* h_ltiles = v.host<int2>(v.o_ltiles);
  int* h_lcount = v.host<int>(v.o_lcount);
  // tiles grouped by order, so each order's kernels run over one contiguous range
  int t = 0;
  for (int j = 0; j < nprobs; ++j) h_probs[j] = slots[j];
  for (int o = 1; o <= kMaxOrder; ++o) {
    h_counts[7 + o] = t;  // offset of order o
    for (int j = 0; j < nprobs; ++j) {
      const Prob& P = *probs_[slots[j]];
      if (P.N != o) continue;
      h_ranges[slots[j]] = make_int2(t, t + (int)(P.S / tile_));
      for (int64_t c = 0; c < P.S; c += tile_) h_tiles[t++] = make_int2(slots[j], (int)c);
    }
    h_counts[2 + o] = t - h_counts[7 + o];
    v.ocap[o] = grow(h_counts[2 + o], v.ocap[o]);
  }
  h_counts[0] = ntiles;
  h_counts[1] = nprobs;
  for (int k = 0; k < kAmgLevels; ++k) {
    int lt = 0;
    for (int sl : slots) {
      const Prob& P = *probs_[sl];
      if (k >= (int)P.amg_rows.size()) continue;
      for (int r0 = 0; r0 < P.amg_rows[k]; r0 += kTile) h_ltiles[(size_t)k * lcap + lt++] = make_int2(sl, r0);
    }
    h_lcount[k] = lt;
  }
  v.levels = levels;
  INSP_HIP(hipMemcpyAsync(v.meta.p, v.h_meta, v.meta_bytes, hipMemcpyHostToDevice, stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
  v.bv.P = d_desc_;
  v.bv.tiles = v.dev<int2>(v.o_tiles);
  v.bv.ntiles = v.dev<int>(v.o_counts);
  v.bv.probs = v.dev<int>(v.o_probs);
  v.bv.ranges = v.dev<int2>(v.o_ranges);
  v.bv.nprobs = v.dev<int>(v.o_counts) + 1;
  v.bv.tile_cap = v.tile_cap;
  v.bv.prob_cap = v.prob_cap;
  v.bv.partial = v.partial.p;
  v.bv.dim = cfg_.dim;
  v.bv.tile = tile_;
  v.bv.ocount = v.dev<int>(v.o_counts) + 2;
  v.bv.ooff = v.dev<int>(v.o_counts) + 7;
  for (int o = 0; o <= kMaxOrder; ++o) v.bv.ocap[o] = v.ocap[o];
  v.bv.ltiles = v.dev<int2>(v.o_ltiles);
  v.bv.lcount = v.dev<int>(v.o_lcount);
  v.bv.ltile_cap = v.ltile_cap;
}

void BatchEngine::drop_graphs() {
  for (auto& kv : graphs_) INSP_HIP(hipGraphExecDestroy(kv.second));
  graphs_.clear();
}

// Resets the convergence mask of these problems to RUNNING; with limits also their
// iteration count, tolerance and iteration limit (one copy and one scatter kernel).
void BatchEngine::write_state(const std::vector<int>& slots, bool with_limits) {
  if (slots.empty()) return;
  INSP_HIP(hipStreamSynchronize(stream_));  // the staging buffer may still feed an earlier copy
  const int n = (int)slots.size();
  for (int j = 0; j < n; ++j) {
    const Prob& P = *probs_[slots[j]];
    double* r = h_stage_ + (size_t)5 * j;
    r[0] = slots[j];
    r[1] = BST_RUNNING;
    r[2] = P.iters;
    r[3] = P.tol;
    r[4] = P.max_iter;
  }
  INSP_HIP(hipMemcpyAsync(d_stage_, h_stage_, sizeof(double) * 5 * n, hipMemcpyHostToDevice, stream_));
  batch::set_state(d_scal_, n, d_stage_, with_limits, stream_);
}

void BatchEngine::read_scal(const std::vector<int>& slots) {
  int hi = 0;
  for (int s : slots) hi = std::max(hi, s + 1);
  if (!hi) return;
  INSP_HIP(hipMemcpyAsync(h_scal_, d_scal_, sizeof(double) * BS_COUNT * hi, hipMemcpyDeviceToHost,
                          stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
}

template <typename T>
void BatchEngine::vcycle(View& v) {
  const BatchView& b = v.bv;
  hipStream_t s = stream_;
  const bool f32 = std::is_same<T, float>::value;
  batch::to_t(b, f32, BV_R, BV_VB, s);
  const int K = cfg_.cheb_order;
  // Chebyshev on order N (level 0) or on order 1 (level 1)
  auto smooth = [&](int level, bool first_zero) {
    const int in = level ? BV_CX : BV_VX, ax = level ? BV_CAP : BV_VAX;
    for (int k = 0; k < K; ++k) {
      if (first_zero && k == 0) {
        batch::cheb<T>(b, level, 0, true, s);
      } else {
        if (level == 0) batch::trace<T>(b, tables_, in, BV_TRT, s);
        batch::apply<T>(b, tables_, level, in, BV_TRT, ax, false, s);
        batch::cheb<T>(b, level, k, false, s);
      }
    }
  };
  for (int it = 0; it < cfg_.pre_it; ++it) smooth(0, it == 0);
  batch::trace<T>(b, tables_, BV_VX, BV_TRT, s);
  batch::apply<T>(b, tables_, 0, BV_VX, BV_TRT, BV_VAX, false, s);
  batch::sub_from_b<T>(b, 0, s);
  batch::restrict_<T>(b, tables_, s);
  if (cfg_.amg) {
    // order 1: smooth, then correct from the vertex space by one AMG V-cycle
    smooth(1, true);
    batch::apply<T>(b, tables_, 1, BV_CX, BV_TRT, BV_CAP, false, s);
    batch::sub_from_b<T>(b, 1, s);
    batch::vertex_restrict<T>(b, s);
    const int L = v.levels, D = cfg_.amg_cheb;
    for (int k = 0; k < L; ++k) {
      batch::amg_dense<T>(b, k, s);
      if (k + 1 == L) break;
      batch::amg_cheb<T>(b, k, 0, true, s);
      for (int j = 1; j < D; ++j) {
        batch::amg_spmv<T>(b, k, false, s);
        batch::amg_cheb<T>(b, k, j, false, s);
      }
      batch::amg_spmv<T>(b, k, true, s);
      batch::amg_restrict<T>(b, k, s);
    }
    for (int k = L - 2; k >= 0; --k) {
      batch::amg_prolong<T>(b, k, s);
      for (int j = 0; j < D; ++j) {
        batch::amg_spmv<T>(b, k, false, s);
        batch::amg_cheb<T>(b, k, j, false, s);
      }
    }
    batch::vertex_prolong<T>(b, s);
    smooth(1, false);
  } else {
    batch::cg_init<T>(b, s);
    batch::finalize(b, 1, BS_CA, s);
    int a = BS_CA, c = BS_CB;
    for (int it = 0; it < cfg_.coarse_iters; ++it) {
      batch::apply<T>(b, tables_, 1, BV_CP, BV_TRT, BV_CAP, true, s);
      batch::finalize(b, 1, BS_CPAP, s);
      batch::cg_xr<T>(b, a, s);
      batch::finalize(b, 1, c, s);
      batch::cg_p<T>(b, c, a, s);
      std::swap(a, c);
    }
  }
  batch::prolong<T>(b, tables_, s);
  for (int it = 0; it < cfg_.post_it; ++it) smooth(0, false);
  batch::from_t(b, f32, BV_VX, BV_Z, s);
  batch::sum_d(b, BV_Z, s);
  batch::finalize(b, 1, BS_MEAN, s);
  batch::sub_mean_d(b, BV_Z, s);
}

void BatchEngine::precondition(View& v) {
  if (cfg_.fp32) vcycle<float>(v);
  else vcycle<double>(v);
}

void BatchEngine::run_iteration(View& v) {
  const BatchView& b = v.bv;
  hipStream_t s = stream_;
  batch::trace<double>(b, tables_, BV_P, BV_TRD, s);
  batch::apply<double>(b, tables_, 0, BV_P, BV_TRD, BV_AP, true, s);
  batch::finalize(b, 1, BS_PAP, s);
  batch::fcg_update_xr(b, s);
  batch::finalize(b, 2, BS_RR, s);  // |r|^2 and r.z_old
  batch::mask(b, s);                 // converged / broken-down problems are frozen from here
  precondition(v);
  batch::dot<double>(b, 0, BV_R, BV_Z, BV_R, BV_Z, 1, s);
  batch::finalize(b, 1, BS_RZN, s);
  batch::fcg_update_p(b, s);
  batch::scal_copy(b, BS_RZ, BS_RZN, s);
}

namespace {
double since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

std::vector<int> BatchEngine::add_many(std::vector<ProblemSpec>&& specs) {
  const int d = cfg_.dim, Nf = d + 1, NK = d == 2 ? 3 : 6;
  auto tp = std::chrono::steady_clock::now();
  std::vector<int> ids;
  std::vector<HostMeshPtr> meshes;
  for (auto& sp : specs) {
    if (!sp.mesh || sp.mesh->dim != d) throw InvalidArgument("problem dimension does not match the batch");
    if (sp.order < 1 || sp.order > kMaxOrder) throw InvalidArgument("order must be in 1..4");
    if (free_slots_.empty()) throw InvalidArgument("batch is full");
    HostMesh& m = *sp.mesh;
    if (m.nbr.empty()) build_topology(m);
    if (m.J.empty()) build_geometry(m);
    auto P = std::make_unique<Prob>();
    P->slot = free_slots_.back();
    P->N = sp.order;
    P->dim = d;
    P->Np = np_of(d, sp.order);
    P->Np1 = d + 1;
    P->Nfp = nfp_of(d, sp.order);
    P->E = m.n_cells;
    P->S = round_up(m.n_cells, kTile);
    P->nbf = m.n_bfaces;
    P->var = !sp.factor.empty();
    P->nullspace = sp.nullspace;
    P->rtol = sp.rtol;
    P->atol = sp.atol;
    P->max_iter = sp.max_iter;
    P->t0 = std::chrono::steady_clock::now();
    P->result.fine_applies = std::max(1, cfg_.max_iter_eig);
    if ((int64_t)sp.rhs.size() != P->E * P->Np) throw InvalidArgument("rhs size mismatch");
    if (P->var && (int64_t)sp.factor.size() != P->E * P->Np) throw InvalidArgument("factor size mismatch");
    if (!sp.bc_types.empty() && (int64_t)sp.bc_types.size() != P->nbf)
      throw InvalidArgument("bc_types size mismatch");
    P->bytes = arena_bytes(d, P->E, P->N, P->var, cfg_.fp32);
    INSP_HIP(hipMalloc(&P->arena, P->bytes));
    INSP_HIP(hipMemset(P->arena, 0, P->bytes));
    INSP_HIP(hipDeviceSynchronize());
    char* base = (char*)P->arena;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      void* p = base + off;
      off += align256(bytes);
      return p;
    };
    const int64_t S = P->S, E = P->E;
    const int Np = P->Np, Np1 = P->Np1, Nfp = P->Nfp;
    const size_t t = cfg_.fp32 ? 4 : 8;
    BatchProb& D = P->desc;
    D.N = P->N; D.Np = Np; D.Np1 = Np1; D.var = P->var; D.nullspace = P->nullspace;
    D.E = E; D.S = S; D.dofs = (double)E * Np;
    double* Kt = (double*)take(NK * S * 8);
    double* J = (double*)take(S * 8);
    int* nbr = (int*)take(Nf * S * 4);
    int* code = (int*)take(Nf * S * 4);
    int* bct = (int*)take((size_t)std::max<int64_t>(E, 1) * Nf * 4);
    double *kF64 = nullptr, *kC64 = nullptr;
    float *kF32 = nullptr, *kC32 = nullptr;
    if (P->var) {
      kF64 = (double*)take(Np * S * 8); kF32 = (float*)take(Np * S * 4);
      kC64 = (double*)take(Np1 * S * 8); kC32 = (float*)take(Np1 * S * 4);
    }
    D.kF[0] = kF64; D.kF[1] = kF32; D.kC[0] = kC64; D.kC[1] = kC32;
    D.v[BV_B] = take(Np * S * 8);
    // everything up to here comes from the host: one image, one copy
    const size_t upload_bytes = off;
    double* tauF = (double*)take(Nf * S * 8);
    double* tauC = (double*)take(Nf * S * 8);
    D.Kt = Kt; D.J = J; D.nbr = nbr; D.code = code; D.bctype = bct; D.tauF = tauF; D.tauC = tauC;
    double* iF64 = (double*)take(Np * S * 8);
    float* iF32 = (float*)take(Np * S * 4);
    double* iC64 = (double*)take(Np1 * S * 8);
    float* iC32 = (float*)take(Np1 * S * 4);
    D.invF[0] = iF64; D.invF[1] = iF32; D.invC[0] = iC64; D.invC[1] = iC32;
    for (int f = BV_B + 1; f <= BV_AP; ++f) D.v[f] = take(Np * S * 8);
    D.v[BV_TRD] = take((size_t)Nf * Nfp * S * 8);
    for (int f = BV_VX; f <= BV_VD; ++f) D.v[f] = take(Np * S * t);
    D.v[BV_TRT] = cfg_.fp32 ? take((size_t)Nf * Nfp * S * t) : D.v[BV_TRD];
    for (int f = BV_CX; f <= BV_CAP; ++f) D.v[f] = take(Np1 * S * t);
    D.scal = d_scal_ + (size_t)P->slot * BS_COUNT;
    if (off > P->bytes) throw std::runtime_error("arena layout overflow");
    // geometry and connectivity, SoA with stride S
    const int nperm = d == 2 ? 2 : 6;
    std::vector<double> kt((size_t)NK * S, 0.0), jj((size_t)S, 1.0);
    std::vector<int> nb((size_t)Nf * S, 0), cd((size_t)Nf * S, 0);
    for (int64_t e = 0; e < E; ++e) {
      for (int k = 0; k < NK; ++k) kt[(size_t)k * S + e] = m.Kt[(size_t)e * NK + k];
      jj[e] = m.J[e];
      for (int f = 0; f < Nf; ++f) {
        const int n = m.nbr[(size_t)e * Nf + f];
        nb[(size_t)f * S + e] = n;
        cd[(size_t)f * S + e] = n >= 0 ? m.nbr_face[(size_t)e * Nf + f] * nperm + m.nbr_perm[(size_t)e * Nf + f] : 0;
      }
    }
    std::vector<char> img(upload_bytes, 0);
    auto put = [&](void* dst, const void* src, size_t nbytes) {
      const size_t at = (size_t)((char*)dst - base);
      if (at + nbytes > upload_bytes) throw std::runtime_error("arena upload layout overflow");
      std::memcpy(img.data() + at, src, nbytes);
    };
    put(Kt, kt.data(), kt.size() * 8);
    put(J, jj.data(), jj.size() * 8);
    put(nbr, nb.data(), nb.size() * 4);
    put(code, cd.data(), cd.size() * 4);
    std::vector<int> bt = sp.bc_types;
    if (bt.empty()) bt.assign((size_t)std::max<int64_t>(P->nbf, 1), 0);
    put(bct, bt.data(), bt.size() * 4);
    const RefElement reN = make_ref_element(d, P->N), re1 = make_ref_element(d, 1);
    if (P->var) {
      std::vector<double> kf((size_t)Np * S, 0.0), kc((size_t)Np1 * S, 0.0);
      const std::vector<double> Ik = interp_matrix(re1, reN);  // Np1 x Np
      for (int64_t e = 0; e < E; ++e) {
        for (int i = 0; i < Np; ++i) kf[(size_t)i * S + e] = sp.factor[(size_t)e * Np + i];
        for (int i = 0; i < Np1; ++i) {
          double a = 0;
          for (int j = 0; j < Np; ++j) a += Ik[(size_t)i * Np + j] * sp.factor[(size_t)e * Np + j];
          kc[(size_t)i * S + e] = a;
        }
      }
      std::vector<float> kff(kf.begin(), kf.end()), kcf(kc.begin(), kc.end());
      put(kF64, kf.data(), kf.size() * 8);
      put(kF32, kff.data(), kff.size() * 4);
      put(kC64, kc.data(), kc.size() * 8);
      put(kC32, kcf.data(), kcf.size() * 4);
    }
    {  // rhs in SoA
      std::vector<double> rb((size_t)Np * S, 0.0);
      for (int64_t e = 0; e < E; ++e)
        for (int i = 0; i < Np; ++i) rb[(size_t)i * S + e] = sp.rhs[(size_t)e * Np + i];
      put(D.v[BV_B], rb.data(), rb.size() * 8);
    }
    INSP_HIP(hipMemcpy(base, img.data(), upload_bytes, hipMemcpyHostToDevice));
    // penalties (finest order in the penalty on both levels) and inverse diagonals
    MeshView mv{d, E, S, Kt, J, nbr, code, bct};
    LevelView lvN{P->N, nullptr, tables_.C64[P->N], tables_.tab[P->N], tauF};
    LevelView lv1{1, nullptr, tables_.C64[1], tables_.tab[1], tauC};
    launch_tau(mv, lvN, kF64, P->N, tauF, stream_);
    launch_tau(mv, lv1, kC64, P->N, tauC, stream_);
    double* tmp = (double*)D.v[BV_R];
    launch_diag(mv, lvN, kF64, tmp, stream_);
    reciprocal<double>(Np, E, S, tmp, iF64, stream_);
    reciprocal<float>(Np, E, S, tmp, iF32, stream_);
    launch_diag(mv, lv1, kC64, tmp, stream_);
    reciprocal<double>(Np1, E, S, tmp, iC64, stream_);
    reciprocal<float>(Np1, E, S, tmp, iC32, stream_);
    fill<double>(Np * S, tmp, 0.0, stream_);
    INSP_HIP(hipStreamSynchronize(stream_));
    if (sp.rhs_nodal) {  // b = J M f
      launch_mass(mv, lvN, (const double*)D.v[BV_B], (double*)D.v[BV_AP], stream_);
      copy<double>((int64_t)Np * S, (const double*)D.v[BV_AP], (double*)D.v[BV_B], stream_);
      fill<double>((int64_t)Np * S, (double*)D.v[BV_AP], 0.0, stream_);
      INSP_HIP(hipStreamSynchronize(stream_));
    }
    bytes_in_use_ += P->bytes;
    const int slot = P->slot;
    free_slots_.pop_back();
    probs_[slot] = std::move(P);
    upload_desc(slot);
    ids.push_back(slot);
    meshes.push_back(sp.mesh);
  }
  if (ids.empty()) return ids;
  setup_s_[0] += since(tp);
  tp = std::chrono::steady_clock::now();
  if (!tuned_) autotune(ids);
  // Chebyshev bounds: power iteration on D^-1 A for all new problems at once, on order N
  // and (AMG coarse solve) on order 1.
  set_view(*work_, ids);
  for (int level = 0; level < (cfg_.amg ? 2 : 1); ++level) {
    const BatchView& b = work_->bv;
    hipStream_t s = stream_;
    batch::rand_d(b, level, BV_R, s);
    for (int it = 0; it < std::max(1, cfg_.max_iter_eig); ++it) {
      if (level == 0) batch::trace<double>(b, tables_, BV_R, BV_TRD, s);
      batch::apply<double>(b, tables_, level, BV_R, BV_TRD, BV_Z, false, s);
      batch::inv_apply_d(b, level, BV_Z, s);
      batch::dot<double>(b, level, BV_R, BV_Z, BV_R, BV_R, 2, s);
      batch::finalize(b, 2, BS_TOP, s);
      batch::dot<double>(b, level, BV_Z, BV_Z, BV_Z, BV_Z, 1, s);
      batch::finalize(b, 1, BS_NRM, s);
      batch::from_t(b, false, BV_Z, BV_R, s);
      batch::scale_d(b, BV_R, BS_NRM, s);
    }
    read_scal(ids);
    for (int slot : ids) {
      Prob& P = *probs_[slot];
      const double* sc = h_scal_ + (size_t)slot * BS_COUNT;
      const double lam = sc[BS_BOT] > 0 ? sc[BS_TOP] / sc[BS_BOT] : 1.0;
      chebyshev_coefficients(cfg_.eig_safety * lam, cfg_.cheb_min_factor, 8,
                             level ? P.desc.amg.oc1 : P.desc.c1, level ? P.desc.amg.oc2 : P.desc.c2);
    }
  }
  setup_s_[1] += since(tp);
  tp = std::chrono::steady_clock::now();
  if (cfg_.amg) setup_amg_all(ids, meshes);
  setup_s_[2] += since(tp);
  tp = std::chrono::steady_clock::now();
  for (int slot : ids)
    INSP_HIP(hipMemcpyAsync(d_desc_ + slot, &probs_[slot]->desc, sizeof(BatchProb), hipMemcpyHostToDevice,
                            stream_));
  // clear the power-iteration and probing scratch (x starts at zero)
  for (int slot : ids) {
    Prob& P = *probs_[slot];
    fill<double>((int64_t)P.Np * P.S, (double*)P.desc.v[BV_R], 0.0, stream_);
    fill<double>((int64_t)P.Np * P.S, (double*)P.desc.v[BV_Z], 0.0, stream_);
  }
  INSP_HIP(hipStreamSynchronize(stream_));
  start(*work_, ids, false);
  setup_s_[3] += since(tp);
  return ids;
}

void BatchEngine::autotune(const std::vector<int>& slots) {
  // Time the order-N smoothing step (trace, apply, Chebyshev update) of the V-cycle precision
  // over these problems for each tile size; the operator dominates a lockstep iteration.
  const int cand[3] = {64, 128, 256};
  TuneResult r{256, {0, 0, 0}};
  hipEvent_t e0, e1;
  INSP_HIP(hipEventCreate(&e0));
  INSP_HIP(hipEventCreate(&e1));
  double best = 1e300;
  for (int q = 0; q < 3; ++q) {
    tile_ = cand[q];
    set_view(*work_, slots);
    const BatchView& b = work_->bv;
    auto step = [&] {
      if (cfg_.fp32) {
        batch::trace<float>(b, tables_, BV_VX, BV_TRT, stream_);
        batch::apply<float>(b, tables_, 0, BV_VX, BV_TRT, BV_VAX, false, stream_);
        batch::cheb<float>(b, 0, 1, false, stream_);
      } else {
        batch::trace<double>(b, tables_, BV_VX, BV_TRT, stream_);
        batch::apply<double>(b, tables_, 0, BV_VX, BV_TRT, BV_VAX, false, stream_);
        batch::cheb<double>(b, 0, 1, false, stream_);
      }
    };
    for (int w = 0; w < 2; ++w) step();
    INSP_HIP(hipEventRecord(e0, stream_));
    for (int w = 0; w < 5; ++w) step();
    INSP_HIP(hipEventRecord(e1, stream_));
    INSP_HIP(hipEventSynchronize(e1));
    float ms = 0;
    INSP_HIP(hipEventElapsedTime(&ms, e0, e1));
    r.ms[q] = ms / 5.0;
    if (r.ms[q] < best) {
      best = r.ms[q];
      r.tile = cand[q];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the timing runs wrote the V-cycle scratch vectors: clear them
  for (int slot : slots) {
    const Prob& P = *probs_[slot];
    const size_t t = cfg_.fp32 ? 4 : 8;
    INSP_HIP(hipMemsetAsync(P.desc.v[BV_VX], 0, (size_t)P.Np * P.S * t, stream_));
    INSP_HIP(hipMemsetAsync(P.desc.v[BV_VAX], 0, (size_t)P.Np * P.S * t, stream_));
    INSP_HIP(hipMemsetAsync(P.desc.v[BV_VD], 0, (size_t)P.Np * P.S * t, stream_));
    INSP_HIP(hipMemsetAsync(P.desc.v[BV_TRT], 0, (size_t)(cfg_.dim + 1) * P.Nfp * P.S * t, stream_));
  }
  INSP_HIP(hipStreamSynchronize(stream_));
  tile_ = r.tile;
  tuned_ = true;
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune.emplace(std::make_pair(cfg_.dim, cfg_.fp32), r);
  tile_ = g_tune[{cfg_.dim, cfg_.fp32}].tile;  // an engine tuned first elsewhere wins
}

BatchEngine::TuneResult BatchEngine::last_tune(int dim, bool fp32) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune.find({dim, fp32});
  return it == g_tune.end() ? TuneResult{0, {0, 0, 0}} : it->second;
}

void BatchEngine::chebyshev_coefficients(double hi, double min_factor, int n, double* c1, double* c2) {
  const double lo = min_factor * hi;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  c1[0] = 0.0;
  c2[0] = 1.0 / theta;
  for (int k = 1; k < n; ++k) {
    const double rho = 1.0 / (2.0 * sigma - rho_old);
    c1[k] = rho * rho_old;
    c2[k] = 2.0 * rho / delta;
    rho_old = rho;
  }
}

size_t BatchEngine::amg_bytes_estimate(int dim, int64_t n_nodes, int64_t cells, bool fp32) {
  // vertex maps + operator (~(7|15) entries per row in 2D|3D) over a hierarchy of about
  // 1.5x the vertex rows, with transfers and four work vectors per row, plus the dense
  // last level.
  const size_t t = fp32 ? 4 : 8, nnz_row = dim == 2 ? 7 : 15;
  const size_t rows = (size_t)(1.5 * (double)n_nodes) + 1;
  size_t b = (size_t)(dim + 1) * cells * 8 + (size_t)n_nodes * 4;   // cv, vlist, vptr
  b += rows * nnz_row * (4 + t) * 3 / 2 + rows * (4 + 5 * t) + 2 * rows * 8 * (4 + t);
  b += (size_t)kAmgMaxCoarse * kAmgMaxCoarse * t + 64 * 1024;
  return b;
}

namespace {
// Runs fn(i) for i < n on up to 16 host threads.
template <typename F>
void parallel_for(size_t n, F&& fn) {
  const size_t hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({n, hw, 16});
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::exception_ptr err;
  std::mutex mu;
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (size_t i; (i = next++) < n;) {
        try {
          fn(i);
        } catch (...) {
          std::lock_guard<std::mutex> g(mu);
          if (!err) err = std::current_exception();
        }
      }
    });
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}
}  // namespace

struct BatchEngine::ProbeScratch {
  // the integer maps of one problem in one buffer (one copy per problem): cv, vptr, vlist, gptr,
  // gcol, colour at off[0..5], each segment sized for the largest problem
  DBuf<int> ints;
  size_t off[7] = {0, 0, 0, 0, 0, 0, 0};
  DBuf<double> val, xv, yv;
};

struct BatchEngine::AmgPrep {
  std::vector<int> vptr, vlist, cv, colour;
  HostCsr g;
  int ncol = 0;
  AmgHierarchy H;
};

void BatchEngine::setup_amg_all(const std::vector<int>& slots, const std::vector<HostMeshPtr>& meshes) {
  std::vector<AmgPrep> prep(slots.size());
  // host: vertex maps, vertex graph, distance-2 colouring
  auto t0 = std::chrono::steady_clock::now();
  parallel_for(slots.size(), [&](size_t q) { amg_prepare(slots[q], *meshes[q], prep[q]); });
  // device: probe each problem's vertex operator, scratch sized for the largest
  ProbeScratch ps;
  size_t ncv = 1, nvp = 1, nvl = 1, ngp = 1, ngc = 1;
  for (const AmgPrep& A : prep) {
    ncv = std::max(ncv, A.cv.size());
    nvp = std::max(nvp, A.vptr.size());
    nvl = std::max(nvl, A.vlist.size());
    ngp = std::max(ngp, A.g.ptr.size());
    ngc = std::max(ngc, A.g.col.size());
  }
  const size_t seg[6] = {ncv, nvp, nvl, ngp, ngc, nvp};
  for (int k = 0; k < 6; ++k) ps.off[k + 1] = ps.off[k] + seg[k];
  ps.ints.alloc(ps.off[6], false);
  ps.val.alloc(ngc, false);
  ps.xv.alloc(nvp, false);
  ps.yv.alloc(nvp, false);
  for (size_t q = 0; q < slots.size(); ++q) amg_probe(slots[q], prep[q], ps);
  // host: hierarchies
  const auto th = std::chrono::steady_clock::now();
  parallel_for(slots.size(), [&](size_t q) { amg_build(slots[q], prep[q]); });
  setup_s_[4] += since(th);
  for (size_t q = 0; q < slots.size(); ++q) amg_upload(slots[q], prep[q]);
  (void)t0;
}

void BatchEngine::amg_prepare(int slot, const HostMesh& m, AmgPrep& A) {
  const Prob& P = *probs_[slot];
  const int d1 = cfg_.dim + 1, nv = (int)m.n_nodes;
  const int64_t E = P.E, S = P.S;
  // vertex maps: order-1 nodes are the cell vertices in cell2nodes order
  A.vptr.assign((size_t)nv + 1, 0);
  A.vlist.resize((size_t)E * d1);
  A.cv.assign((size_t)d1 * S, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int i = 0; i < d1; ++i) ++A.vptr[(size_t)m.cell2nodes[e * d1 + i] + 1];
  for (int v = 0; v < nv; ++v) A.vptr[v + 1] += A.vptr[v];
  std::vector<int> pos(A.vptr.begin(), A.vptr.end() - 1);
  for (int64_t e = 0; e < E; ++e)
    for (int i = 0; i < d1; ++i) {
      const int v = m.cell2nodes[e * d1 + i];
      A.vlist[pos[v]++] = (int)(i * S + e);
      A.cv[(size_t)i * S + e] = v;
    }
  A.g = vertex_graph(m);
  A.colour = distance2_colouring(A.g, &A.ncol);
}

void BatchEngine::amg_probe(int slot, AmgPrep& A, ProbeScratch& ps) {
  // A_v = P_v^T A_1 P_v by distance-2 colour classes (double, this problem alone)
  Prob& P = *probs_[slot];
  const int d1 = cfg_.dim + 1, nv = (int)A.vptr.size() - 1;
  const int64_t E = P.E, S = P.S;
  const std::vector<int>* src[6] = {&A.cv, &A.vptr, &A.vlist, &A.g.ptr, &A.g.col, &A.colour};
  std::vector<int> h(ps.off[6], 0);
  for (int k = 0; k < 6; ++k) {
    if (src[k]->size() > ps.off[k + 1] - ps.off[k]) throw std::runtime_error("probe scratch overflow");
    std::copy(src[k]->begin(), src[k]->end(), h.begin() + (std::ptrdiff_t)ps.off[k]);
  }
  INSP_HIP(hipMemcpyAsync(ps.ints.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
  struct {
    int* p;
  } d_cv{ps.ints.p + ps.off[0]}, d_vptr{ps.ints.p + ps.off[1]}, d_vlist{ps.ints.p + ps.off[2]},
      gptr{ps.ints.p + ps.off[3]}, gcol{ps.ints.p + ps.off[4]}, dcol{ps.ints.p + ps.off[5]};
  DBuf<double>& val = ps.val;
  DBuf<double>& xv = ps.xv;
  DBuf<double>& yv = ps.yv;
  set_view(*work_, {slot});
  double* u = (double*)P.desc.v[BV_R];
  double* y = (double*)P.desc.v[BV_Z];
  for (int c = 0; c < A.ncol; ++c) {
    amg::probe_fill(nv, dcol.p, c, xv.p, stream_);
    amg::vertex_prolong<double>(d1, E, S, d_cv.p, xv.p, u, false, stream_);
    batch::apply<double>(work_->bv, tables_, 1, BV_R, BV_TRD, BV_Z, false, stream_);
    amg::vertex_restrict<double>(nv, d_vptr.p, d_vlist.p, y, yv.p, stream_);
    amg::probe_extract(nv, gptr.p, gcol.p, dcol.p, c, yv.p, val.p, stream_);
  }
  INSP_HIP(hipMemcpyAsync(A.g.val.data(), val.p, A.g.val.size() * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
}

void BatchEngine::amg_build(int slot, AmgPrep& A) {
  const Prob& P = *probs_[slot];
  {  // the hierarchy is built from the symmetric part
    HostCsr gt = csr_transpose(A.g);
    for (size_t k = 0; k < A.g.val.size(); ++k) A.g.val[k] = 0.5 * (A.g.val[k] + gt.val[k]);
  }
  AmgOptions opt;
  opt.max_coarse = kAmgMaxCoarse;
  opt.max_levels = kAmgLevels;
  opt.cheb_order = cfg_.amg_cheb;
  A.H = amg_setup(std::move(A.g), opt, P.nullspace);
  if (A.H.n_coarse > kAmgMaxCoarse) throw std::runtime_error("AMG coarsening stalled above the dense size");
}

void BatchEngine::amg_upload(int slot, AmgPrep& A) {
  Prob& P = *probs_[slot];
  const size_t t = cfg_.fp32 ? 4 : 8;
  const int nv = (int)A.vptr.size() - 1;
  const AmgHierarchy& H = A.H;
  const std::vector<int>& cv = A.cv;
  const std::vector<int>& vptr = A.vptr;
  const std::vector<int>& vlist = A.vlist;
  // one device allocation for the maps and the hierarchy
  auto a = [](size_t bytes) { return align256(bytes); };
  size_t bytes = a(cv.size() * 4) + a(vptr.size() * 4) + a(vlist.size() * 4);
  for (size_t k = 0; k < H.levels.size(); ++k) {
    const AmgHostLevel& h = H.levels[k];
    bytes += a((h.A.ptr.size()) * 4) + a(h.A.col.size() * 4) + a(h.A.val.size() * t) + 5 * a((size_t)h.A.n * t);
    if (k + 1 < H.levels.size())
      bytes += a(h.P.ptr.size() * 4) + a(h.P.col.size() * 4) + a(h.P.val.size() * t) + a(h.R.ptr.size() * 4) +
               a(h.R.col.size() * 4) + a(h.R.val.size() * t);
  }
  bytes += a((size_t)H.n_coarse * H.n_coarse * t);
  INSP_HIP(hipMalloc(&P.amg_arena, bytes));
  P.amg_bytes = bytes;
  bytes_in_use_ += bytes;
  // the arena is laid out in a host image (zero where the device fills it) and sent in one copy
  std::vector<char> img(bytes, 0);
  char* base = (char*)P.amg_arena;
  size_t off = 0;
  auto put = [&](const void* h, size_t nbytes) -> void* {
    if (off + a(nbytes) > bytes) throw std::runtime_error("AMG arena layout overflow");
    void* d = base + off;
    if (h && nbytes) std::memcpy(img.data() + off, h, nbytes);
    off += a(nbytes);
    return d;
  };
  auto put_t = [&](const std::vector<double>& v) -> void* {
    if (cfg_.fp32) {
      std::vector<float> f(v.begin(), v.end());
      return put(f.data(), f.size() * 4);
    }
    return put(v.data(), v.size() * 8);
  };
  BatchAmg& D = P.desc.amg;
  D.cv = (const int*)put(cv.data(), cv.size() * 4);
  D.vptr = (const int*)put(vptr.data(), vptr.size() * 4);
  D.vlist = (const int*)put(vlist.data(), vlist.size() * 4);
  D.nv = nv;
  D.nlev = (int)H.levels.size();
  P.amg_rows.clear();
  for (size_t k = 0; k < H.levels.size(); ++k) {
    const AmgHostLevel& h = H.levels[k];
    BatchAmgLevel& L = D.lev[k];
    L.n = h.A.n;
    L.m = k + 1 < H.levels.size() ? h.P.m : 0;
    L.Aptr = (const int*)put(h.A.ptr.data(), h.A.ptr.size() * 4);
    L.Acol = (const int*)put(h.A.col.data(), h.A.col.size() * 4);
    L.Aval = put_t(h.A.val);
    std::vector<double> inv((size_t)h.A.n);
    for (int i = 0; i < h.A.n; ++i) inv[i] = h.diag[i] != 0 ? 1.0 / h.diag[i] : 0.0;
    L.inv = put_t(inv);
    L.x = put(nullptr, (size_t)h.A.n * t);
    L.b = put(nullptr, (size_t)h.A.n * t);
    L.r = put(nullptr, (size_t)h.A.n * t);
    L.d = put(nullptr, (size_t)h.A.n * t);
    if (k + 1 < H.levels.size()) {
      L.Pptr = (const int*)put(h.P.ptr.data(), h.P.ptr.size() * 4);
      L.Pcol = (const int*)put(h.P.col.data(), h.P.col.size() * 4);
      L.Pval = put_t(h.P.val);
      L.Rptr = (const int*)put(h.R.ptr.data(), h.R.ptr.size() * 4);
      L.Rcol = (const int*)put(h.R.col.data(), h.R.col.size() * 4);
      L.Rval = put_t(h.R.val);
    }
    chebyshev_coefficients(1.1 * h.lmax, 1.0 / 20.0, 4, L.ch1, L.ch2);
    P.amg_rows.push_back(h.A.n);
  }
  D.nc = H.n_coarse;
  D.cinv = put_t(H.coarse_inv);
  INSP_HIP(hipMemcpy(base, img.data(), bytes, hipMemcpyHostToDevice));
}

// r = b - A x, convergence test, then z = M r, p = z for problems not yet converged.
void BatchEngine::start(View& v, const std::vector<int>& slots, bool restart) {
  if (slots.empty()) return;
  set_view(v, slots);
  const BatchView& b = v.bv;
  hipStream_t s = stream_;
  write_state(slots, false);  // unfreeze: every kernel below works on these problems
  if (!restart) {
    batch::sum_d(b, BV_B, s);
    batch::finalize(b, 1, BS_MEAN, s);
    batch::sub_mean_d(b, BV_B, s);
    batch::dot<double>(b, 0, BV_B, BV_B, BV_B, BV_B, 1, s);
    batch::finalize(b, 1, BS_BB, s);
  }
  batch::trace<double>(b, tables_, BV_X, BV_TRD, s);
  batch::apply<double>(b, tables_, 0, BV_X, BV_TRD, BV_AP, false, s);
  batch::residual(b, s);
  batch::finalize(b, 1, BS_RR, s);
  read_scal(slots);
  std::vector<int> go;
  for (int slot : slots) {
    Prob& P = *probs_[slot];
    const double* sc = h_scal_ + (size_t)slot * BS_COUNT;
    ++P.result.fine_applies;
    if (!restart) {
      P.bnorm = std::sqrt(sc[BS_BB]);
      P.tol = std::max(P.rtol * P.bnorm, P.atol);
    }
    const double rr = sc[BS_RR];
    if (!std::isfinite(rr) || !std::isfinite(P.bnorm)) {
      finish(slot, ProbStatus::BREAKDOWN);
    } else if (std::sqrt(rr) <= P.tol) {
      P.result.rel_residual = P.bnorm > 0 ? std::sqrt(rr) / P.bnorm : 0.0;
      finish(slot, ProbStatus::CONVERGED);
    } else if (P.iters >= P.max_iter) {
      P.result.rel_residual = std::sqrt(rr) / P.bnorm;
      finish(slot, ProbStatus::NOT_CONVERGED);
    } else {
      go.push_back(slot);
    }
  }
  if (go.empty()) return;
  set_view(v, go);
  write_state(go, true);  // tolerance, iteration count and limit of the device mask
  precondition(v);
  batch::dot<double>(v.bv, 0, BV_R, BV_Z, BV_R, BV_Z, 1, s);
  batch::finalize(v.bv, 1, BS_RZ, s);
  batch::from_t(v.bv, false, BV_Z, BV_P, s);
  INSP_HIP(hipStreamSynchronize(stream_));
  for (int slot : go) {
    running_slots_.push_back(slot);
    ++probs_[slot]->result.vcycles;
  }
  std::sort(running_slots_.begin(), running_slots_.end());
  graph_tile_cap_ = -2;  // running set changed: refresh the view before the next replay
}

void BatchEngine::finish(int slot, ProbStatus st) {
  Prob& P = *probs_[slot];
  P.result.status = st;
  P.result.iterations = P.iters;
  P.result.bnorm = P.bnorm;
  P.result.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - P.t0).count();
  P.result.x.assign((size_t)P.E * P.Np, 0.0);
  DBuf<double> tmp;
  tmp.alloc((size_t)P.E * P.Np, false);
  aos_from_soa(P.Np, P.E, P.S, (const double*)P.desc.v[BV_X], tmp.p, stream_);
  INSP_HIP(hipMemcpyAsync(P.result.x.data(), tmp.p, sizeof(double) * P.E * P.Np,
                          hipMemcpyDeviceToHost, stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
  P.done = true;
  finished_.push_back(slot);
}

int BatchEngine::iterate(int k) {
  const int n = iterate_launch(k);
  if (n) iterate_collect();
  return n;
}

int BatchEngine::iterate_launch(int k) {
  launched_ = 0;
  if (running_slots_.empty() || k < 1) return 0;
  const bool refresh = graph_tile_cap_ == -2 || !run_->meta.p;
  if (refresh) {
    const int old_t = run_->tile_cap, old_p = run_->prob_cap, old_l = run_->ltile_cap,
              old_lev = run_->levels;
    int old_o[5];
    std::copy(run_->ocap, run_->ocap + 5, old_o);
    set_view(*run_, running_slots_);
    // the graphs bake in the grid sizes and the number of AMG levels they walk
    if (run_->tile_cap != old_t || run_->prob_cap != old_p || run_->ltile_cap != old_l ||
        run_->levels != old_lev || !std::equal(old_o, old_o + 5, run_->ocap))
      drop_graphs();
    graph_tile_cap_ = run_->tile_cap;
    graph_prob_cap_ = run_->prob_cap;
  }
  if (cfg_.use_graph) {
    auto it = graphs_.find(k);
    if (it == graphs_.end()) {
      hipGraph_t g;
      hipGraphExec_t exec;
      INSP_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < k; ++i) run_iteration(*run_);
      INSP_HIP(hipStreamEndCapture(stream_, &g));
      INSP_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
      INSP_HIP(hipGraphDestroy(g));
      it = graphs_.emplace(k, exec).first;
    }
    INSP_HIP(hipGraphLaunch(it->second, stream_));
  } else {
    for (int i = 0; i < k; ++i) run_iteration(*run_);
  }
  iters_run_ += k;
  // the problems' state comes back with the iterations: the only host round trip of the k
  // iterations, waited for in iterate_collect
  int hi = 0;
  for (int s : running_slots_) hi = std::max(hi, s + 1);
  INSP_HIP(hipMemcpyAsync(h_scal_, d_scal_, sizeof(double) * BS_COUNT * hi, hipMemcpyDeviceToHost, stream_));
  launched_ = k;
  return k;
}

void BatchEngine::iterate_collect() {
  if (!launched_) return;
  launched_ = 0;
  INSP_HIP(hipStreamSynchronize(stream_));
  std::vector<int> keep, verify;
  for (int slot : running_slots_) {
    Prob& P = *probs_[slot];
    const double* sc = h_scal_ + (size_t)slot * BS_COUNT;
    const int its = (int)sc[BS_ITER];
    const int state = (int)sc[BS_STATE];
    const int ran = its - P.iters;
    P.iters = its;
    P.result.fine_applies += ran;
    // the iteration that froze a problem stops before its preconditioner
    P.result.vcycles += ran - (state != BST_RUNNING ? 1 : 0);
    if (state == BST_BREAKDOWN) {
      finish(slot, ProbStatus::BREAKDOWN);
    } else if (state == BST_TOL || state == BST_MAXIT) {
      verify.push_back(slot);
    } else {
      keep.push_back(slot);
    }
  }
  if (keep.size() != running_slots_.size()) {
    running_slots_ = keep;
    graph_tile_cap_ = -2;
  }
  if (!verify.empty()) start(*work_, verify, true);
}

int BatchEngine::running() const { return (int)running_slots_.size(); }

std::vector<int> BatchEngine::take_finished() {
  std::vector<int> f;
  f.swap(finished_);
  return f;
}

const ProbResult& BatchEngine::result(int id) const {
  if (id < 0 || id >= max_slots_ || !probs_[id]) throw InvalidArgument("unknown problem id");
  return probs_[id]->result;
}

void BatchEngine::release(int id) {
  if (id < 0 || id >= max_slots_ || !probs_[id]) throw InvalidArgument("unknown problem id");
  if (!probs_[id]->done) throw InvalidArgument("problem is still running");
  INSP_HIP(hipStreamSynchronize(stream_));
  (void)hipFree(probs_[id]->arena);
  if (probs_[id]->amg_arena) (void)hipFree(probs_[id]->amg_arena);
  bytes_in_use_ -= probs_[id]->bytes + probs_[id]->amg_bytes;
  probs_[id].reset();
  free_slots_.push_back(id);
}

}  // namespace insp
