#include "solver.h"

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <cmath>
#include <cstring>

namespace insp {

namespace {
// Device scalar slots.
enum : int {
  S_RZ = 0, S_PAP = 1, S_RR = 2, S_RZN = 3, S_BNUM = 4,
  S_CG_A = 8, S_CG_PAP = 9, S_CG_B = 10, S_CG_ASAVE = 11,
  S_EIG_TOP = 12, S_EIG_BOT = 13, S_EIG_NRM = 14,
  S_MEAN = 16, S_BB = 17,
  S_COUNT = 32
};
static_assert(S_BNUM == S_RZN + 1, "dot_rz_flex writes r.z and r.(z - zold) to adjacent slots");

int64_t round_up(int64_t v, int64_t m) { return ((v + m - 1) / m) * m; }
}  // namespace

template <typename T>
struct PoissonSolver::Work {
  DBuf<T> x, b, ax, d;     // V-cycle vectors
  DBuf<T> x2;              // smoothing levels: the next iterate of a fused Chebyshev step
  DBuf<T> z, p, p2, Ap;    // coarse CG vectors (coarsest level only; p, p2: two directions)
  DBuf<T> tr;              // face traces
  DBuf<T> inv;             // inverse diagonal
  DBuf<T> k;               // factor (empty = constant)
  DBuf<T> mm;              // per-node mass-term factor (empty = constant or none)
  DBuf<T> C;               // operator constants
  DBuf<T> frag;            // matrix-core fragments of the element pass (empty: vector pass)
  DBuf<T> P, Pt;           // transfer to/from the next coarser level
};

struct PoissonSolver::Level {
  int N = 0;
  OpLayout L{};
  RefElement re;
  DBuf<double> C64;
  DBuf<int> tab;
  DBuf<double> tauS, diag, Ik;
  DBuf<double> tau_cells;  // 3D order 1: tauS cell by cell, for the coarse CG's apply
  double lmax = 0;
  Work<double> w64;
  Work<float> w32;
  template <typename T>
  Work<T>& work();
};
template <>
PoissonSolver::Work<double>& PoissonSolver::Level::work<double>() { return w64; }
template <>
PoissonSolver::Work<float>& PoissonSolver::Level::work<float>() { return w32; }

// Device copy of one AMG level in one precision.
template <typename T>
struct AmgLevelBufs {
  DBuf<T> Aval, Pval, Rval, inv;   // operator, transfers, inverse diagonal
  DBuf<T> x, b, r, d;              // iterate, rhs, residual / A x, Chebyshev direction
  DBuf<T> b0, acc;                 // vertex level only: rhs and sum over repeated cycles
};

struct PoissonSolver::Amg {
  // vertex space
  int nv = 0;
  DBuf<int> vptr, vlist;           // vertex -> SoA entries of the order-1 nodes at it
  DBuf<int> cv;                    // order-1 SoA entry -> vertex
  HostCsr Av;                      // vertex operator (host copy)
  // hierarchy
  struct Lev {
    int n = 0, m = 0;              // rows, rows of the next level
    int lanesA = 1, lanesP = 1, lanesR = 1;  // lanes per row of each product
    DBuf<int> Aptr, Acol, Pptr, Pcol, Rptr, Rcol;
    AmgLevelBufs<double> f64;
    AmgLevelBufs<float> f32;
    double lmax = 1.0;
    template <typename T>
    AmgLevelBufs<T>& bufs();
  };
  std::vector<std::unique_ptr<Lev>> lev;
  int nc = 0;
  DBuf<double> cinv64;
  DBuf<float> cinv32;
  int cheb_order = 2;
  double cheb_ratio = 20.0;        // lower Chebyshev bound = upper / ratio
};
template <>
AmgLevelBufs<double>& PoissonSolver::Amg::Lev::bufs<double>() { return f64; }
template <>
AmgLevelBufs<float>& PoissonSolver::Amg::Lev::bufs<float>() { return f32; }

PoissonSolver::PoissonSolver(HostMesh mesh, int order, hipStream_t stream)
    : mesh_(std::move(mesh)), N_(order), stream_(stream) {
  if (mesh_.dim != 2 && mesh_.dim != 3) throw InvalidArgument("mesh dimension must be 2 or 3");
  if (order < 1 || order > kMaxOrder) throw InvalidArgument("order must be in 1..4");
  if (mesh_.n_cells < 1) throw InvalidArgument("mesh has no cells");
  if (mesh_.nbr.empty()) build_topology(mesh_);
  if (mesh_.J.empty()) build_geometry(mesh_);
  fine_re_ = make_ref_element(mesh_.dim, order);
  if (!stream_) {
    INSP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  const int64_t E = mesh_.n_cells;
  // SoA row pitch: a multiple of 64 cells, and an odd one, so that the rows of one cell (its
  // values, metric, traces) do not all fall on the same cache sets as they would with a pitch of
  // 2^k times a small factor (64^3 x 6 cells = 3 x 2^19)
  stride_ = round_up(E, 64);
  if (((stride_ / 64) & 1) == 0) stride_ += 64;
  const int d = mesh_.dim, Nf = d + 1, NK = d == 2 ? 3 : 6;
  const int nperm = d == 2 ? 2 : 6;
  std::vector<double> kt((size_t)NK * stride_, 0.0), jj((size_t)stride_, 0.0);
  std::vector<int> nb((size_t)Nf * stride_, 0), cd((size_t)Nf * stride_, 0);
  for (int64_t e = 0; e < E; ++e) {
    for (int k = 0; k < NK; ++k) kt[(size_t)k * stride_ + e] = mesh_.Kt[(size_t)e * NK + k];
    jj[e] = mesh_.J[e];
    for (int f = 0; f < Nf; ++f) {
      const int n = mesh_.nbr[(size_t)e * Nf + f];
      nb[(size_t)f * stride_ + e] = n;
      cd[(size_t)f * stride_ + e] =
          n >= 0 ? mesh_.nbr_face[(size_t)e * Nf + f] * nperm + mesh_.nbr_perm[(size_t)e * Nf + f] : 0;
    }
  }
  Kt_.upload(kt.data(), kt.size());
  if (d == 3) Ka_.upload(mesh_.Kt.data(), (size_t)E * NK);
  J_.upload(jj.data(), jj.size());
  nbr_.upload(nb.data(), nb.size());
  code_.upload(cd.data(), cd.size());
  if (E < (int64_t(1) << 26) && Nf * nperm <= 32) {
    for (size_t i = 0; i < nb.size(); ++i)
      if (nb[i] >= 0) nb[i] = (nb[i] << 5) | cd[i];
    nbc_.upload(nb.data(), nb.size());
    if (d == 3) {
      std::vector<int> nc((size_t)E * Nf);
      for (int64_t e = 0; e < E; ++e)
        for (int f = 0; f < Nf; ++f) nc[(size_t)e * Nf + f] = nb[(size_t)f * stride_ + e];
      nbc_cells_.upload(nc.data(), nc.size());
    }
  }
  std::vector<int> bt((size_t)std::max<int64_t>(mesh_.n_bfaces, 1), 0);
  bctype_.upload(bt.data(), bt.size());
  scal_.alloc(S_COUNT);
  partial_.alloc((size_t)std::max<int64_t>(4 * kRedBlocks, apply_blocks(E)));
  INSP_HIP(hipHostMalloc((void**)&host_scal_, S_COUNT * sizeof(double), hipHostMallocDefault));
}

PoissonSolver::~PoissonSolver() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& pr : ev_trace_) { ev_pool_.push_back(pr.first); ev_pool_.push_back(pr.second); }
  for (auto& pr : ev_apply_) { ev_pool_.push_back(pr.first); ev_pool_.push_back(pr.second); }
  for (auto& pr : ev_coarse_) { ev_pool_.push_back(pr.first); ev_pool_.push_back(pr.second); }
  std::sort(ev_pool_.begin(), ev_pool_.end());
  ev_pool_.erase(std::unique(ev_pool_.begin(), ev_pool_.end()), ev_pool_.end());
  for (auto ev : ev_pool_) (void)hipEventDestroy(ev);
  if (host_scal_) (void)hipHostFree(host_scal_);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

MeshView PoissonSolver::mesh_view() const {
  MeshView m;
  m.dim = mesh_.dim;
  m.E = mesh_.n_cells;
  m.stride = stride_;
  m.Kt = Kt_.p;
  m.J = J_.p;
  m.nbr = nbr_.p;
  m.code = code_.p;
  m.bctype = bctype_.p;
  m.nbc = nbc_.empty() ? nullptr : nbc_.p;
  m.Ka = Ka_.empty() ? nullptr : Ka_.p;
  m.nbc_cells = nbc_cells_.empty() ? nullptr : nbc_cells_.p;
  return m;
}

const double* PoissonSolver::fine_factor() const {
  return var_factor_ ? levels_[0]->w64.k.p : nullptr;
}
const double* PoissonSolver::fine_mm() const {
  return var_mm_ ? levels_[0]->w64.mm.p : nullptr;
}
int PoissonSolver::fine_face_nodes() const {
  return levels_[0]->L.Nf * levels_[0]->L.Nfp;
}

bool PoissonSolver::matrix_pass(int l) const {
  return l >= 0 && l < (int)levels_.size() && !levels_[l]->w32.frag.empty();
}

LevelView PoissonSolver::level_view(int l, bool fp32) const {
  const Level& lv = *levels_[l];
  LevelView v;
  v.N = lv.N;
  v.C = fp32 ? (const void*)lv.w32.C.p : (const void*)lv.w64.C.p;
  v.C64 = lv.C64.p;
  v.tab = lv.tab.p;
  v.tauS = lv.tauS.p;
  v.tau_cells = lv.tau_cells.empty() ? nullptr : lv.tau_cells.p;
  v.frag = fp32 ? (const void*)lv.w32.frag.p : (const void*)lv.w64.frag.p;
  return v;
}

void PoissonSolver::set_bc_types(const int* types, int64_t n) {
  if (n != mesh_.n_bfaces) throw InvalidArgument("bc_types length must equal the number of bfaces");
  if (n == 0) return;
  if (!types) throw InvalidArgument("bc_types is null");
  bctype_.upload(types, (size_t)n);
  initialised_ = false;
}

void PoissonSolver::set_factor(const double* k, Layout layout) {
  if (!k) {
    var_factor_ = false;
    factor_.release();
  } else {
    factor_.alloc((size_t)np() * stride_);
    import_vec(k, factor_.p, layout, np());
    var_factor_ = true;
  }
  initialised_ = false;
}

const OiRule& PoissonSolver::oi_rule() {
  if (!oi_rule_) oi_rule_ = std::make_unique<OiRule>(make_oi_rule(mesh_.dim, N_));
  return *oi_rule_;
}

void PoissonSolver::ensure_oi_tables() {
  if (!oi_C64_.empty()) return;
  const OiRule& q = oi_rule();
  const std::vector<double> C = make_oi_tables(fine_re_, q);
  if (!oi_shape_matches(mesh_.dim, N_, q.nv, q.ns, (int)C.size()))
    throw std::logic_error("over-integration rule and kernels disagree");
  oi_C64_.upload(C.data(), C.size());
  std::vector<float> Cf(C.begin(), C.end());
  oi_C32_.upload(Cf.data(), Cf.size());
  oi_qmap_.upload(q.qmap.data(), q.qmap.size());
}

// One part of the coefficient at the points (volume: nv rows, faces: Nf*ns rows), kept in both
// precisions; null clears it. Over-integration is on when both parts are set.
void PoissonSolver::set_oi_part(const double* k, Layout layout, bool surface) {
  DBuf<double>& d = surface ? oi_ks64_ : oi_kv64_;
  DBuf<float>& f = surface ? oi_ks32_ : oi_kv32_;
  if (!k) {
    d.release();
    f.release();
  } else {
    ensure_oi_tables();
    const OiRule& q = oi_rule();
    const int rows = surface ? (mesh_.dim + 1) * q.ns : q.nv;
    d.alloc((size_t)rows * stride_);
    f.alloc((size_t)rows * stride_);
    import_vec(k, d.p, layout, rows);
    convert<double, float>((int64_t)rows * stride_, d.p, f.p, stream_);
    INSP_HIP(hipStreamSynchronize(stream_));
  }
  oi_ = !oi_kv64_.empty() && !oi_ks64_.empty();
  initialised_ = false;
}

void PoissonSolver::set_factor_oi(const double* kv, Layout layout) { set_oi_part(kv, layout, false); }
void PoissonSolver::set_factor_surf_oi(const double* ks, Layout layout) { set_oi_part(ks, layout, true); }

template <>
OiView<double> PoissonSolver::oi_view<double>() {
  const int64_t n = (int64_t)(mesh_.dim + 1) * oi_rule().ns * stride_;
  if (oi_tr64_.empty()) oi_tr64_.alloc((size_t)2 * n);
  return OiView<double>{oi_C64_.p, oi_qmap_.p, oi_kv64_.p, oi_ks64_.p, oi_tr64_.p, oi_tr64_.p + n};
}
template <>
OiView<float> PoissonSolver::oi_view<float>() {
  const int64_t n = (int64_t)(mesh_.dim + 1) * oi_rule().ns * stride_;
  if (oi_tr32_.empty()) oi_tr32_.alloc((size_t)2 * n);
  return OiView<float>{oi_C32_.p, oi_qmap_.p, oi_kv32_.p, oi_ks32_.p, oi_tr32_.p, oi_tr32_.p + n};
}

void PoissonSolver::set_mm_factor(const double* mm, Layout layout) {
  if (!mm) {
    var_mm_ = false;
    mm_.release();
  } else {
    mm_.alloc((size_t)np() * stride_);
    import_vec(mm, mm_.p, layout, np());
    var_mm_ = true;
  }
  initialised_ = false;
}

void PoissonSolver::set_mm_const(double mm) {
  if (!std::isfinite(mm)) throw InvalidArgument("mass factor must be finite");
  mm_const_ = mm;
  initialised_ = false;
}

void PoissonSolver::set_inv_mass_recp_factor(const double* f, Layout layout) {
  if (!f) {
    inv_mass_recp_.release();
    return;
  }
  inv_mass_recp_.alloc((size_t)np() * stride_);
  import_vec(f, inv_mass_recp_.p, layout, np());
}

void PoissonSolver::grad_central(const double* p, double* const g[3], const double* const u[3], double dt,
                                 Layout layout) {
  if (!initialised_) init();
  const int d = mesh_.dim, Np = np(), Nf = d + 1, Nfp = nfp();
  if (!p) throw InvalidArgument("null pressure");
  for (int x = 0; x < d; ++x)
    if (!g[x]) throw InvalidArgument("null gradient component");
  if (u && !std::isfinite(dt)) throw InvalidArgument("dt must be finite");
  const int64_t E = mesh_.n_cells;
  if (grad_xi_.empty()) {
    std::vector<double> G((size_t)E * d * d), soa((size_t)d * d * stride_, 0.0);
    mesh_inverse_jacobians(mesh_, G.data());
    for (int64_t e = 0; e < E; ++e)
      for (int k = 0; k < d * d; ++k) soa[(size_t)k * stride_ + e] = G[(size_t)e * d * d + k];
    grad_xi_.upload(soa.data(), soa.size());
    // LIFT = M^-1 EMAT, EMAT[i][(f, j)] = face mass of face f at volume node i
    const std::vector<double> Mi = mat_inverse(fine_re_.M, Np);
    const int NF = Nf * Nfp;
    std::vector<double> lift((size_t)Np * NF, 0.0);
    for (int i = 0; i < Np; ++i)
      for (int f = 0; f < Nf; ++f)
        for (int j = 0; j < Nfp; ++j) {
          double a = 0;
          for (int l = 0; l < Nfp; ++l)
            a += Mi[(size_t)i * Np + fine_re_.fmask[f * Nfp + l]] *
                 fine_re_.Mface[((size_t)f * Nfp + l) * Nfp + j];
          lift[(size_t)i * NF + f * Nfp + j] = a;
        }
    lift_.upload(lift.data(), lift.size());
  }
  const size_t n = (size_t)Np * stride_;
  DBuf<double> pd;
  pd.alloc(n);
  import_vec(p, pd.p, layout, Np);
  DBuf<double> ud[3], od[3];
  const double* up[3] = {nullptr, nullptr, nullptr};
  double* op[3] = {nullptr, nullptr, nullptr};
  for (int x = 0; x < d; ++x) {
    od[x].alloc(n);
    op[x] = od[x].p;
    if (u) {
      if (!u[x]) throw InvalidArgument("null velocity component");
      ud[x].alloc(n);
      import_vec(u[x], ud[x].p, layout, Np);
      up[x] = ud[x].p;
    }
  }
  const Level& l0 = *levels_[0];
  launch_grad_central(d, N_, E, stride_, grad_xi_.p, nbr_.p, code_.p, l0.C64.p, l0.tab.p, lift_.p, pd.p,
                      dt, up, op, stream_);
  for (int x = 0; x < d; ++x) export_vec(od[x].p, g[x], layout, Np);
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::set_bcs(const double* bc, Layout layout) {
  if (!bc) {
    have_bc_ = false;
    return;
  }
  const size_t n = (size_t)mesh_.n_bfaces * nfp();
  if (n == 0) {
    have_bc_ = false;
    return;
  }
  bc_.alloc(n);
  if (layout == Layout::HOST_AOS)
    INSP_HIP(hipMemcpyAsync(bc_.p, bc, n * sizeof(double), hipMemcpyHostToDevice, stream_));
  else
    INSP_HIP(hipMemcpyAsync(bc_.p, bc, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
  have_bc_ = true;
}

void PoissonSolver::set_tol_and_iter(double rtol, double atol, int max_iter) {
  if (!(rtol >= 0) || !(atol >= 0) || max_iter < 0) throw InvalidArgument("invalid tolerances");
  rtol_ = rtol;
  atol_ = atol;
  max_iter_ = max_iter;
}

void make_op_tables(const RefElement& re, std::vector<double>& C, std::vector<int>& tab) {
  const int d = re.dim, N = re.N;
  const OpLayout L = op_layout(d, N);
  const int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf;
  const FmaskTable fm = make_fmask(d, N);
  for (int i = 0; i < Nf * Nfp; ++i)
    if (fm.v[i] != re.fmask[i]) throw std::runtime_error("face mask tables disagree");
  C.assign((size_t)L.size, 0.0);
  for (int s = 0; s < d; ++s)
    for (int j = 0; j < Np; ++j)
      for (int m = 0; m < Np; ++m)
        C[L.off_D + (s * Np + j) * Np + m] = re.D[s][(size_t)j * Np + m];
  for (int r = 0; r < d; ++r)
    for (int j = 0; j < Np; ++j)
      for (int i = 0; i < Np; ++i) {
        double a = 0;
        for (int k = 0; k < Np; ++k) a += re.D[r][(size_t)k * Np + i] * re.M[(size_t)k * Np + j];
        C[L.off_Bt + (r * Np + j) * Np + i] = a;
      }
  for (int f = 0; f < Nf; ++f)
    for (int j = 0; j < Nfp; ++j)
      for (int i = 0; i < Nfp; ++i)
        C[L.off_MfT + (f * Nfp + j) * Nfp + i] = re.Mface[((size_t)f * Nfp + i) * Nfp + j];
  for (int f = 0; f < Nf; ++f)
    for (int r = 0; r < d; ++r)
      for (int j = 0; j < Nfp; ++j)
        for (int i = 0; i < Np; ++i) {
          double a = 0;
          for (int l = 0; l < Nfp; ++l)
            a += re.D[r][(size_t)re.fmask[f * Nfp + l] * Np + i] *
                 re.Mface[((size_t)f * Nfp + l) * Nfp + j];
          C[L.off_Wt + ((f * d + r) * Nfp + j) * Np + i] = a;
        }
  for (int i = 0; i < Np * Np; ++i) C[L.off_M + i] = re.M[i];
  if (N == 1) {
    // the order-1 kernels take the basis derivatives, the mass column sums and the face mass as
    // compile-time constants: they must be the tables' values
    auto near = [](double a, double b) { return std::fabs(a - b) <= 1e-13; };
    bool ok = true;
    for (int r = 0; r < d; ++r)
      for (int k = 0; k < Np; ++k)
        for (int i = 0; i < Np; ++i) ok = ok && near(re.D[r][(size_t)k * Np + i], p1_dphi(r, i));
    for (int j = 0; j < Np; ++j) {
      double a = 0;
      for (int k = 0; k < Np; ++k) a += re.M[(size_t)k * Np + j];
      ok = ok && near(a, p1_mass_colsum(d));
    }
    for (int f = 0; f < Nf; ++f)
      for (int i = 0; i < Nfp; ++i)
        for (int j = 0; j < Nfp; ++j)
          ok = ok && near(re.Mface[((size_t)f * Nfp + i) * Nfp + j], p1_face_mass_c(d) * (i == j ? 2 : 1));
    if (!ok) throw std::logic_error("order-1 reference tables differ from the P1 constants");
  }
  tab.assign((size_t)L.tab_size, -1);
  for (int i = 0; i < Nf * Nfp; ++i) tab[L.tab_fmask + i] = re.fmask[i];
  for (int f = 0; f < Nf; ++f)
    for (int fn = 0; fn < Nf; ++fn)
      for (int p = 0; p < L.nperm; ++p)
        for (int j = 0; j < Nfp; ++j) {
          const size_t idx = (((size_t)f * Nf + fn) * L.nperm + p) * Nfp + j;
          const int q = re.nbr_face_node[idx];
          tab[L.tab_nbr + idx] = (re.fmask[fn * Nfp + q] << 6) | (fn * Nfp + q);
        }
  for (int f = 0; f < Nf; ++f)
    for (int l = 0; l < Nfp; ++l) tab[L.tab_finv + f * Np + re.fmask[f * Nfp + l]] = l;
}


void PoissonSolver::build_levels() {
  std::vector<int> orders;
  if (type_ == SolverType::PMULTIGRID) {
    orders = pmg_.orders;
    if (orders.empty())
      for (int n = N_; n >= 1; --n) orders.push_back(n);
    if (orders.front() != N_) throw InvalidArgument("p-multigrid orders must start at the solver order");
    for (size_t i = 1; i < orders.size(); ++i)
      if (orders[i] >= orders[i - 1] || orders[i] < 1)
        throw InvalidArgument("p-multigrid orders must decrease strictly");
  } else {
    orders = {N_};
  }
  const int d = mesh_.dim;
  levels_.clear();
  for (int N : orders) {
    auto lv = std::make_unique<Level>();
    lv->N = N;
    lv->L = op_layout(d, N);
    lv->re = make_ref_element(d, N);
    const OpLayout& L = lv->L;
    const RefElement& re = lv->re;
    const int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf;
    std::vector<double> C;
    std::vector<int> tab;
    make_op_tables(re, C, tab);
    lv->C64.upload(C.data(), C.size());
    lv->tab.upload(tab.data(), tab.size());
    lv->tauS.alloc((size_t)Nf * stride_);
    lv->diag.alloc((size_t)Np * stride_);
    // operator constants in both precisions
    lv->w64.C.upload(C.data(), C.size());
    std::vector<float> Cf(C.begin(), C.end());
    lv->w32.C.upload(Cf.data(), Cf.size());
    const bool matrix = epass_ == ElementPass::MATRIX || (epass_ == ElementPass::AUTO && d == 3 && N == 4);
    const std::vector<double> fr = matrix ? build_mfma_fragments(d, N, C, tab) : std::vector<double>();
    if (!fr.empty()) {
      std::vector<float> frf(fr.begin(), fr.end());
      lv->w32.frag.upload(frf.data(), frf.size());
    }
    levels_.push_back(std::move(lv));
  }
  for (size_t l = 0; l + 1 < levels_.size(); ++l) {
    Level& F = *levels_[l];
    Level& Cc = *levels_[l + 1];
    const int npf = F.L.Np, npc = Cc.L.Np;
    std::vector<double> P = interp_matrix(F.re, Cc.re);  // npf x npc
    std::vector<double> Pt = mat_transpose(P, npf, npc);
    std::vector<double> Ik = interp_matrix(Cc.re, F.re);  // npc x npf
    F.w64.P.upload(P.data(), P.size());
    F.w64.Pt.upload(Pt.data(), Pt.size());
    std::vector<float> Pf(P.begin(), P.end()), Ptf(Pt.begin(), Pt.end());
    F.w32.P.upload(Pf.data(), Pf.size());
    F.w32.Pt.upload(Ptf.data(), Ptf.size());
    F.Ik.upload(Ik.data(), Ik.size());
  }
}

void PoissonSolver::ensure_factor_levels() {
  const int64_t E = mesh_.n_cells;
  // per-node mass-term factor on every level (interpolated like the coefficient)
  for (size_t l = 0; l < levels_.size(); ++l) {
    Level& lv = *levels_[l];
    if (!var_mm_) {
      lv.w64.mm.release();
      lv.w32.mm.release();
      continue;
    }
    const size_t n = (size_t)lv.L.Np * stride_;
    lv.w64.mm.alloc(n);
    if (l == 0) {
      copy<double>((int64_t)n, mm_.p, lv.w64.mm.p, stream_);
    } else {
      Level& F = *levels_[l - 1];
      launch_elem_matvec<double>(mesh_.dim, lv.N, F.N, E, stride_, F.Ik.p, F.w64.mm.p, lv.w64.mm.p,
                                 false, stream_);
    }
    lv.w32.mm.alloc(n);
    convert<double, float>((int64_t)n, lv.w64.mm.p, lv.w32.mm.p, stream_);
  }
  for (size_t l = 0; l < levels_.size(); ++l) {
    Level& lv = *levels_[l];
    if (!var_factor_) {
      lv.w64.k.release();
      lv.w32.k.release();
      continue;
    }
    const size_t n = (size_t)lv.L.Np * stride_;
    lv.w64.k.alloc(n);
    if (l == 0) {
      copy<double>((int64_t)n, factor_.p, lv.w64.k.p, stream_);
    } else {
      Level& F = *levels_[l - 1];
      launch_elem_matvec<double>(mesh_.dim, lv.N, F.N, E, stride_, F.Ik.p, F.w64.k.p, lv.w64.k.p,
                                 false, stream_);
    }
    if (prec_ == Precision::MIXED || l == 0) {
      lv.w32.k.alloc(n);
      convert<double, float>((int64_t)n, lv.w64.k.p, lv.w32.k.p, stream_);
    }
  }
}

void PoissonSolver::compute_diagonals() {
  const MeshView mv = mesh_view();
  for (size_t l = 0; l < levels_.size(); ++l) {
    Level& lv = *levels_[l];
    const int Np = lv.L.Np;
    if (oi_ && l == 0)
      launch_diag_oi(mv, level_view(0, false), oi_view<double>(), lv.diag.p, stream_);
    else
      launch_diag(mv, level_view((int)l, false), var_factor_ ? lv.w64.k.p : nullptr, lv.diag.p, stream_);
    if (has_mass_term())
      launch_mass_diag(mv, level_view((int)l, false), Np, var_mm_ ? lv.w64.mm.p : nullptr, mm_const_,
                       lv.diag.p, stream_);
    lv.w64.inv.alloc((size_t)Np * stride_);
    reciprocal<double>(Np, mesh_.n_cells, stride_, lv.diag.p, lv.w64.inv.p, stream_);
    lv.w32.inv.alloc((size_t)Np * stride_);
    reciprocal<float>(Np, mesh_.n_cells, stride_, lv.diag.p, lv.w32.inv.p, stream_);
  }
}

double PoissonSolver::read_scalar(int slot) {
  INSP_HIP(hipMemcpyAsync(host_scal_, scal_.p, S_COUNT * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
  return host_scal_[slot];
}

void PoissonSolver::estimate_eigenvalues() {
  // Power iteration on D^-1 A with a Rayleigh quotient (the reference's
  // set_rand_vec / vec_norm / vec_normalise / rayleigh_quotient kernels).
  RedScratch rs{partial_.p};
  // With the AMG coarse solve the coarsest order is smoothed as well.
  const size_t smoothed = pmg_.coarse == CoarseType::AMG ? levels_.size() : levels_.size() - 1;
  for (size_t l = 0; l < smoothed; ++l) {
    Level& lv = *levels_[l];
    Work<double>& w = lv.w64;
    const int Np = lv.L.Np;
    const int64_t n = (int64_t)Np * stride_;
    DBuf<double> v, av, tr_save;
    v.alloc(n);
    av.alloc(n);
    if (w.tr.empty()) w.tr.alloc((size_t)lv.L.Nf * lv.L.Nfp * stride_);
    rand_fill(Np, mesh_.n_cells, stride_, 0x5eedULL + 977ULL * lv.N, v.p, stream_);
    double lam = 0;
    for (int it = 0; it < std::max(1, pmg_.max_iter_eig); ++it) {
      apply_level<double>((int)l, v.p, av.p);
      jacobi<double>(n, av.p, w.inv.p, av.p, stream_);
      dot<double>(n, v.p, av.p, v.p, v.p, 2, scal_.p, S_EIG_TOP, rs, stream_);
      dot<double>(n, av.p, av.p, nullptr, nullptr, 1, scal_.p, S_EIG_NRM, rs, stream_);
      read_scalar(0);
      lam = host_scal_[S_EIG_TOP] / host_scal_[S_EIG_BOT];
      const double nrm = std::sqrt(host_scal_[S_EIG_NRM]);
      if (!(nrm > 0) || !std::isfinite(nrm)) break;
      axpby<double>(n, 1.0 / nrm, av.p, 0.0, v.p, stream_);
    }
    INSP_HIP(hipStreamSynchronize(stream_));  // v, av are released below
    lv.lmax = lam;
  }
}

void PoissonSolver::init() {
  build_levels();
  ensure_factor_levels();
  const MeshView mv = mesh_view();
  // Every p-multigrid level uses the finest order in the penalty, as the reference's
  // kernels use the program-wide DG_ORDER (fvmf_2d_calc_tau_faces.h:9); this keeps the
  // rediscretised coarse operators equal to the Galerkin products P^T A P.
  for (size_t l = 0; l < levels_.size(); ++l)
    launch_tau(mv, level_view((int)l, false), var_factor_ ? levels_[l]->w64.k.p : nullptr, N_,
               levels_[l]->tauS.p, stream_);
  for (size_t l = 0; l < levels_.size(); ++l)
    if (mesh_.dim == 3 && levels_[l]->N == 1 && !nbc_cells_.empty()) {
      Level& lv = *levels_[l];
      lv.tau_cells.alloc((size_t)lv.L.Nf * stride_, false);
      rows_to_cells<double>(mesh_.n_cells, stride_, lv.L.Nf, lv.tauS.p, lv.tau_cells.p, stream_);
    }
  compute_diagonals();
  // work vectors
  const bool mixed = prec_ == Precision::MIXED;
  for (size_t l = 0; l < levels_.size(); ++l) {
    Level& lv = *levels_[l];
    const size_t n = (size_t)lv.L.Np * stride_;
    const size_t nt = (size_t)lv.L.Nf * lv.L.Nfp * stride_;
    auto setup = [&](auto& w) {
      w.tr.alloc(nt);
      if (type_ == SolverType::PMULTIGRID) {
        w.x.alloc(n); w.b.alloc(n); w.ax.alloc(n); w.d.alloc(n);
        if (l + 1 < levels_.size() || pmg_.coarse == CoarseType::AMG) w.x2.alloc(n);
        if (l + 1 == levels_.size() && pmg_.coarse == CoarseType::JACOBI_CG) {
          w.z.alloc(n); w.p.alloc(n); w.p2.alloc(n); w.Ap.alloc(n);
        }
      }
    };
    setup(lv.w64);
    if (mixed) setup(lv.w32);
  }
  const size_t n0 = (size_t)np() * stride_;
  b_.alloc(n0); x_.alloc(n0); r_.alloc(n0); z_.alloc(n0); zold_.alloc(n0); p_.alloc(n0); Ap_.alloc(n0);
  if (mixed_refinement()) {
    z32_.alloc(n0); zold32_.alloc(n0); p32_.alloc(n0); Ap32_.alloc(n0);
  } else {
    z32_.release(); zold32_.release(); p32_.release(); Ap32_.release();
  }
  if (type_ == SolverType::PMULTIGRID) estimate_eigenvalues();
  amg_.reset();
  if (type_ == SolverType::PMULTIGRID && pmg_.coarse == CoarseType::AMG) build_coarse_space();
  if (type_ == SolverType::BLOCK_JACOBI_PCG) build_block_jacobi();
  else blk_.release();
  if (type_ == SolverType::INV_MASS_PCG) {
    // M^-1 = V V^T for the nodal basis; computed here from M itself
    const std::vector<double> Mi = mat_inverse(fine_re_.M, np());
    minv_.upload(Mi.data(), Mi.size());
  } else {
    minv_.release();
  }
  INSP_HIP(hipStreamSynchronize(stream_));
  (void)noalloc_scratch();  // this device's copy of the allocation-free apply's scratch
  initialised_ = true;
}

hipEvent_t PoissonSolver::ev_take() {
  if (!ev_pool_.empty()) {
    hipEvent_t e = ev_pool_.back();
    ev_pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  INSP_HIP(hipEventCreate(&e));
  return e;
}

template <typename T>
void PoissonSolver::apply_level(int l, const T* u, T* out, double* partial, T* tr, bool cells) {
  Level& lv = *levels_[l];
  Work<T>& w = lv.work<T>();
  const bool own_tr = tr == nullptr;  // a caller-given trace buffer: no allocation, no events
  if (own_tr) {
    if (w.tr.empty()) w.tr.alloc((size_t)lv.L.Nf * lv.L.Nfp * stride_);
    tr = w.tr.p;
  }
  const MeshView mv = mesh_view();
  const LevelView v = level_view(l, std::is_same<T, float>::value);
  const T* k = var_factor_ ? w.k.p : nullptr;
  const bool prof = own_tr && profiling_ && l == 0;
  const bool prof_coarse = own_tr && profiling_ && l > 0 && l + 1 == (int)levels_.size() &&
                           (coarse_sample_++ & 3) == 0;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  if (prof) {
    e0 = ev_take(); e1 = ev_take(); e2 = ev_take();
    INSP_HIP(hipEventRecord(e0, stream_));
  } else if (prof_coarse) {
    e0 = ev_take(); e2 = ev_take();
    INSP_HIP(hipEventRecord(e0, stream_));
  }
  if (cells) {
    // u and out cell by cell (the coarse Jacobi-CG's copies; 3D order 1, no mass term)
    if (prof) INSP_HIP(hipEventRecord(e1, stream_));
    launch_apply_p1_cells<T>(mv, v, u, k, out, stream_, partial);
  } else if (oi_ && l == 0) {
    if (prof) INSP_HIP(hipEventRecord(e1, stream_));
    launch_apply_oi<T>(mv, v, oi_view<T>(), u, out, stream_, partial);
  } else {
    launch_trace<T>(mv, v, u, k, tr, stream_);
    if (prof) INSP_HIP(hipEventRecord(e1, stream_));
    launch_apply<T>(mv, v, u, tr, k, out, stream_, partial);
  }
  if (has_mass_term()) {
    if (cells) throw std::logic_error("cell-by-cell apply with a mass term");
    launch_mass_add<T>(mv, v, lv.L.Np, var_mm_ ? w.mm.p : nullptr, (T)mm_const_, u, out, stream_,
                       partial);
  }
  if (prof) {
    INSP_HIP(hipEventRecord(e2, stream_));
    ev_trace_.push_back({e0, e1});
    ev_apply_.push_back({e1, e2});
  } else if (prof_coarse) {
    INSP_HIP(hipEventRecord(e2, stream_));
    ev_coarse_.push_back({e0, e2});
  }
}

CoarseTiming PoissonSolver::take_coarse_timing() {
  INSP_HIP(hipStreamSynchronize(stream_));
  CoarseTiming t;
  for (auto& pr : ev_coarse_) {
    float ms = 0;
    INSP_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    t.apply_ms += ms;
    t.launches++;
    ev_pool_.push_back(pr.first);
    ev_pool_.push_back(pr.second);
  }
  ev_coarse_.clear();
  return t;
}

KernelTiming PoissonSolver::take_timing() {
  INSP_HIP(hipStreamSynchronize(stream_));
  for (auto& pr : ev_trace_) {
    float ms = 0;
    INSP_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    timing_.trace_ms += ms;
    timing_.trace_launches++;
  }
  for (auto& pr : ev_apply_) {
    float ms = 0;
    INSP_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
    timing_.apply_ms += ms;
    timing_.apply_launches++;
  }
  std::vector<hipEvent_t> used;
  for (auto& pr : ev_trace_) { used.push_back(pr.first); used.push_back(pr.second); }
  for (auto& pr : ev_apply_) used.push_back(pr.second);
  ev_pool_.insert(ev_pool_.end(), used.begin(), used.end());
  ev_trace_.clear();
  ev_apply_.clear();
  KernelTiming t = timing_;
  timing_ = KernelTiming{};
  return t;
}

template <typename T>
void PoissonSolver::chebyshev(int l, const T* f, T* x, bool zero_init) {
  Level& lv = *levels_[l];
  Work<T>& w = lv.work<T>();
  const int64_t n = (int64_t)lv.L.Np * stride_;
  const int K = l < (int)pmg_.cheb_orders.size() ? std::max(1, pmg_.cheb_orders[l]) : 3;
  const double hi = pmg_.eig_safety * lv.lmax;
  const double lo = pmg_.cheb_min_factor * hi;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
  const double sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  // A step after an apply: fused into the element pass where the level has the vector pass and
  // no mass term, else apply + stream. Where x is the level's own iterate the step is the
  // three-term form x_next = x + c1 (x - x_prev) + c2 D^-1 (f - A x): x_next overwrites x_prev in
  // x2, then x and x2 trade places, so x2 holds the previous iterate at every step
  // (d = x - x_prev is not stored).
  const MeshView mv = mesh_view();
  const LevelView v = level_view(l, std::is_same<T, float>::value);
  const bool three = x == w.x.p && !w.x2.empty();
  const bool fuse = three && !has_mass_term() && !(oi_ && l == 0) &&
                    !(v.frag && mfma_apply_supported(mv.dim, v.N));
  bool prev_zero = false;  // x_prev = 0: the first iterate came from x = 0 by the stream step
  auto step = [&](T c1, T c2) {
    ++smoother_applies_;
    if (fuse) {
      if (w.tr.empty()) w.tr.alloc((size_t)lv.L.Nf * lv.L.Nfp * stride_);
      const T* k = var_factor_ ? w.k.p : nullptr;
      launch_trace<T>(mv, v, x, k, w.tr.p, stream_);
      const T* prev = prev_zero ? nullptr : w.x2.p;
      prev_zero = false;
      if (!launch_apply_cheb<T>(mv, v, x, w.tr.p, k, f, w.inv.p, prev, w.x2.p, c1, c2, stream_))
        throw std::logic_error("fused Chebyshev step not available on this level");
      std::swap(w.x, w.x2);
      x = w.x.p;
      return;
    }
    apply_level<T>(l, x, w.ax.p);
    if (three) {
      const T* prev = prev_zero ? nullptr : w.x2.p;
      prev_zero = false;
      cheb_step3<T>(n, f, w.ax.p, w.inv.p, prev, x, w.x2.p, c1, c2, stream_);
      std::swap(w.x, w.x2);
      x = w.x.p;
      return;
    }
    cheb_step<T>(n, f, w.ax.p, w.inv.p, w.d.p, x, c1, c2, false, stream_);
  };
  if (zero_init) {
    // the first iterate from x = 0; the three-term steps need no d
    cheb_step<T>(n, f, nullptr, w.inv.p, three ? nullptr : w.d.p, x, (T)0, (T)(1.0 / theta), true,
                 stream_);
    prev_zero = true;
  } else
    step((T)0, (T)(1.0 / theta));
  for (int k = 2; k <= K; ++k) {
    const double rho = 1.0 / (2.0 * sigma - rho_old);
    step((T)(rho * rho_old), (T)(2.0 * rho / delta));
    rho_old = rho;
  }
}

template <typename T>
void PoissonSolver::coarse_cg(int l) {
  Level& lv = *levels_[l];
  Work<T>& w = lv.work<T>();
  const int64_t n = (int64_t)lv.L.Np * stride_;
  RedScratch rs{partial_.p};
  // Jacobi-CG carried in the preconditioned residual z = D^-1 r alone: per iteration the
  // apply (p.Ap fused), z -= alpha D^-1 Ap with r.z = sum z^2 D, then the direction update. x is
  // brought up to date every second iteration, from the two last directions (two buffers): the
  // same operations in the same order as one x update per iteration, with 4.5 instead of 5
  // vector passes per iteration.
  //
  // In 3D at order 1 the iteration runs on cell-by-cell copies (v[e*4 + i]): the apply then reads a
  // cell's four values, its own and each neighbour's, with one or two 16-byte loads instead of four
  // rows. b and D^-1 are copied in (into the level's ax and d, unused by this solve) and x is
  // copied out at the end; the vector kernels are the same over the flat range.
  const bool cells = mesh_.dim == 3 && lv.N == 1 && !has_mass_term() && !(oi_ && l == 0) &&
                     !matrix_pass(l) && !Ka_.empty() &&
                     !nbc_cells_.empty() && !lv.tau_cells.empty();
  const T* bb = w.b.p;
  const T* inv = w.inv.p;
  if (cells) {
    rows_to_cells<T>(mesh_.n_cells, stride_, lv.L.Np, w.b.p, w.ax.p, stream_);
    rows_to_cells<T>(mesh_.n_cells, stride_, lv.L.Np, w.inv.p, w.d.p, stream_);
    bb = w.ax.p;
    inv = w.d.p;
  }
  fill<T>(n, w.x.p, (T)0, stream_);
  jacobi<T>(n, bb, inv, w.z.p, stream_);
  copy<T>(n, w.z.p, w.p.p, stream_);
  dot<T>(n, bb, w.z.p, nullptr, nullptr, 1, scal_.p, S_CG_A, rs, stream_);
  int a = S_CG_A, b = S_CG_B;
  T* pc = w.p.p;   // the current direction
  T* pq = w.p2.p;  // the previous one (odd iterations) or the buffer for the next one
  for (int it = 0; it < pmg_.coarse_iters; ++it) {
    if (std::is_same<T, float>::value && matrix_pass(l) && !(oi_ && l == 0)) {
      // the matrix-core element pass has no fused dot
      apply_level<T>(l, pc, w.Ap.p);
      dot<T>(n, pc, w.Ap.p, nullptr, nullptr, 1, scal_.p, S_CG_PAP, rs, stream_);
    } else {
      apply_level<T>(l, pc, w.Ap.p, partial_.p, nullptr, cells);  // p.Ap fused into the pass
      finalize(apply_blocks(mesh_.n_cells), 1, partial_.p, scal_.p, S_CG_PAP, stream_);
    }
    ++smoother_applies_;
    cg_update_z<T>(n, w.z.p, w.Ap.p, inv, scal_.p, a, S_CG_PAP, b, rs, stream_);
    const bool last = it + 1 == pmg_.coarse_iters;
    if (it % 2 == 1) {
      // x += alpha_{k-1} p_{k-1} + alpha_k p_k; the next direction overwrites p_{k-1}
      cg_update_x2p<T>(n, w.x.p, pq, pc, w.z.p, scal_.p, S_CG_ASAVE, a, S_CG_PAP, b, last, stream_);
      std::swap(pc, pq);
    } else if (last) {
      cg_update_xp<T>(n, w.x.p, pc, w.z.p, scal_.p, a, S_CG_PAP, b, true, stream_);
    } else {
      // the next direction into the other buffer; alpha_k kept for the x update
      cg_update_p<T>(n, pq, pc, w.z.p, scal_.p, a, S_CG_PAP, b, S_CG_ASAVE, stream_);
      std::swap(pc, pq);
    }
    std::swap(a, b);
  }
  if (cells) {
    // x back to rows through ax, then x and ax trade places
    cells_to_rows<T>(mesh_.n_cells, stride_, lv.L.Np, w.x.p, w.ax.p, stream_);
    std::swap(w.x, w.ax);
  }
}

void PoissonSolver::build_block_jacobi() {
  // Greedy colouring of the face-neighbour graph: no two neighbours share a colour, so a
  // unit vector on every cell of one colour returns each of those cells' own block column.
  const int64_t E = mesh_.n_cells;
  const int Nf = mesh_.dim + 1, Np = np();
  std::vector<int> colour((size_t)E, -1);
  int ncol = 0;
  for (int64_t e = 0; e < E; ++e) {
    unsigned used = 0;
    for (int f = 0; f < Nf; ++f) {
      const int n = mesh_.nbr[(size_t)e * Nf + f];
      if (n >= 0 && colour[n] >= 0) used |= 1u << colour[n];
    }
    int c = 0;
    while (used & (1u << c)) ++c;
    colour[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  DBuf<int> dcol;
  dcol.upload(colour.data(), colour.size());
  blk_.alloc((size_t)Np * Np * stride_);
  DBuf<double> u, y;
  u.alloc((size_t)Np * stride_);
  y.alloc((size_t)Np * stride_);
  for (int c = 0; c < ncol; ++c)
    for (int j = 0; j < Np; ++j) {
      block_probe_unit(Np, E, stride_, dcol.p, c, j, u.p, stream_);
      apply_level<double>(0, u.p, y.p);
      block_extract_column(Np, E, stride_, dcol.p, c, j, y.p, blk_.p, stream_);
    }
  block_invert(Np, E, stride_, blk_.p, stream_);
  INSP_HIP(hipStreamSynchronize(stream_));  // u, y, dcol are released on return
}

void PoissonSolver::inv_block_diag(double* out) {
  if (!initialised_) init();
  if (type_ != SolverType::BLOCK_JACOBI_PCG || blk_.empty())
    throw InvalidArgument("inverse blocks exist only for the block-Jacobi solver type");
  const int Np = np();
  const int64_t E = mesh_.n_cells;
  std::vector<double> h((size_t)Np * Np * stride_);
  INSP_HIP(hipMemcpyAsync(h.data(), blk_.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
  for (int64_t e = 0; e < E; ++e)
    for (int k = 0; k < Np * Np; ++k) out[(size_t)e * Np * Np + k] = h[(size_t)k * stride_ + e];
}

void PoissonSolver::build_coarse_space() {
  const int l = (int)levels_.size() - 1;
  Level& lv = *levels_[l];
  if (lv.N != 1) throw InvalidArgument("the AMG coarse solve needs order 1 as the coarsest p-multigrid order");
  auto A = std::make_unique<Amg>();
  const int d1 = mesh_.dim + 1;
  const int64_t E = mesh_.n_cells;
  const int nv = (int)mesh_.n_nodes;
  A->nv = nv;
  // order-1 nodes are the cell vertices in cell2nodes order
  std::vector<int> vptr((size_t)nv + 1, 0), vlist((size_t)E * d1), cv((size_t)d1 * stride_, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int i = 0; i < d1; ++i) ++vptr[(size_t)mesh_.cell2nodes[e * d1 + i] + 1];
  for (int v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  {
    std::vector<int> pos(vptr.begin(), vptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int i = 0; i < d1; ++i) {
        const int v = mesh_.cell2nodes[e * d1 + i];
        vlist[pos[v]++] = (int)(i * stride_ + e);
        cv[(size_t)i * stride_ + e] = v;
      }
  }
  A->vptr.upload(vptr.data(), vptr.size());
  A->vlist.upload(vlist.data(), vlist.size());
  A->cv.upload(cv.data(), cv.size());
  // assemble A_v = P_v^T A_1 P_v by probing with distance-2 colour classes
  HostCsr g = vertex_graph(mesh_);
  int ncol = 0;
  const std::vector<int> colour = distance2_colouring(g, &ncol);
  DBuf<int> gptr, gcol, dcol;
  gptr.upload(g.ptr.data(), g.ptr.size());
  gcol.upload(g.col.data(), g.col.size());
  dcol.upload(colour.data(), colour.size());
  DBuf<double> val, xv, yv, u, y;
  val.alloc(g.col.size());
  xv.alloc((size_t)nv);
  yv.alloc((size_t)nv);
  u.alloc((size_t)d1 * stride_);
  y.alloc((size_t)d1 * stride_);
  for (int c = 0; c < ncol; ++c) {
    amg::probe_fill(nv, dcol.p, c, xv.p, stream_);
    amg::vertex_prolong<double>(d1, E, stride_, A->cv.p, xv.p, u.p, false, stream_);
    apply_level<double>(l, u.p, y.p);
    amg::vertex_restrict<double>(nv, A->vptr.p, A->vlist.p, y.p, yv.p, stream_);
    amg::probe_extract(nv, gptr.p, gcol.p, dcol.p, c, yv.p, val.p, stream_);
  }
  g.val.resize(g.col.size());
  INSP_HIP(hipMemcpyAsync(g.val.data(), val.p, g.val.size() * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
  INSP_HIP(hipStreamSynchronize(stream_));
  // A_v is the exact Galerkin product. With a variable coefficient the SIPG operator takes
  // k from the owner side in its symmetric face term and is not exactly symmetric; the AMG
  // hierarchy is built from the symmetric part (the pattern is symmetric).
  A->Av = g;
  {
    HostCsr gt = csr_transpose(g);
    for (size_t k = 0; k < g.val.size(); ++k) g.val[k] = 0.5 * (g.val[k] + gt.val[k]);
  }
  AmgOptions opt;
  opt.cheb_order = A->cheb_order;
  bool singular = nullspace_;
  AmgHierarchy H = amg_setup(std::move(g), opt, singular);
  const bool mixed = prec_ == Precision::MIXED;
  for (size_t k = 0; k < H.levels.size(); ++k) {
    const AmgHostLevel& h = H.levels[k];
    auto L = std::make_unique<Amg::Lev>();
    L->n = h.A.n;
    L->m = h.P.m;
    L->lmax = h.lmax;
    L->Aptr.upload(h.A.ptr.data(), h.A.ptr.size());
    L->Acol.upload(h.A.col.data(), h.A.col.size());
    L->lanesA = lanes_per_row((int64_t)h.A.col.size(), h.A.n);
    std::vector<double> inv((size_t)h.A.n);
    for (int i = 0; i < h.A.n; ++i) inv[i] = h.diag[i] != 0 ? 1.0 / h.diag[i] : 0.0;
    const bool has_next = k + 1 < H.levels.size();
    if (has_next) {
      L->Pptr.upload(h.P.ptr.data(), h.P.ptr.size());
      L->Pcol.upload(h.P.col.data(), h.P.col.size());
      L->Rptr.upload(h.R.ptr.data(), h.R.ptr.size());
      L->Rcol.upload(h.R.col.data(), h.R.col.size());
      L->lanesP = lanes_per_row((int64_t)h.P.col.size(), h.P.n);
      L->lanesR = lanes_per_row((int64_t)h.R.col.size(), h.R.n);
    }
    auto fill_bufs = [&](auto& B, auto tag) {
      using T = decltype(tag);
      std::vector<T> a(h.A.val.begin(), h.A.val.end()), iv(inv.begin(), inv.end());
      B.Aval.upload(a.data(), a.size());
      B.inv.upload(iv.data(), iv.size());
      if (has_next) {
        std::vector<T> p(h.P.val.begin(), h.P.val.end()), r(h.R.val.begin(), h.R.val.end());
        B.Pval.upload(p.data(), p.size());
        B.Rval.upload(r.data(), r.size());
      }
      B.x.alloc((size_t)std::max(h.A.n, 1));
      B.b.alloc((size_t)std::max(h.A.n, 1));
      B.r.alloc((size_t)std::max(h.A.n, 1));
      B.d.alloc((size_t)std::max(h.A.n, 1));
      if (k == 0 && pmg_.amg_cycles > 1) {
        B.b0.alloc((size_t)std::max(h.A.n, 1));
        B.acc.alloc((size_t)std::max(h.A.n, 1));
      }
    };
    fill_bufs(L->f64, double{});
    if (mixed) fill_bufs(L->f32, float{});
    A->lev.push_back(std::move(L));
  }
  A->nc = H.n_coarse;
  A->cinv64.upload(H.coarse_inv.data(), H.coarse_inv.size());
  if (mixed) {
    std::vector<float> ci(H.coarse_inv.begin(), H.coarse_inv.end());
    A->cinv32.upload(ci.data(), ci.size());
  }
  amg_ = std::move(A);
}

template <typename T>
void PoissonSolver::amg_vcycle(int k) {
  Amg& A = *amg_;
  Amg::Lev& L = *A.lev[k];
  AmgLevelBufs<T>& B = L.bufs<T>();
  const DevCsr<T> Am{L.n, L.n, L.Aptr.p, L.Acol.p, B.Aval.p, L.lanesA};
  if (k + 1 == (int)A.lev.size()) {
    const T* M = std::is_same<T, float>::value ? (const T*)(const void*)A.cinv32.p
                                               : (const T*)(const void*)A.cinv64.p;
    amg::dense_matvec<T>(A.nc, M, B.b.p, B.x.p, stream_);
    return;
  }
  const double hi = 1.1 * L.lmax, lo = hi / A.cheb_ratio;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  auto smooth = [&](bool zero_init) {
    double rho_old = 1.0 / sigma;
    if (zero_init) {
      cheb_step<T>(L.n, B.b.p, nullptr, B.inv.p, B.d.p, B.x.p, (T)0, (T)(1.0 / theta), true, stream_);
    } else {
      amg::spmv<T>(Am, B.x.p, B.r.p, nullptr, 0, stream_);
      cheb_step<T>(L.n, B.b.p, B.r.p, B.inv.p, B.d.p, B.x.p, (T)0, (T)(1.0 / theta), false, stream_);
    }
    for (int j = 2; j <= A.cheb_order; ++j) {
      amg::spmv<T>(Am, B.x.p, B.r.p, nullptr, 0, stream_);
      const double rho = 1.0 / (2.0 * sigma - rho_old);
      cheb_step<T>(L.n, B.b.p, B.r.p, B.inv.p, B.d.p, B.x.p, (T)(rho * rho_old), (T)(2.0 * rho / delta),
                   false, stream_);
      rho_old = rho;
    }
  };
  smooth(true);
  amg::spmv<T>(Am, B.x.p, B.r.p, B.b.p, 2, stream_);  // r = b - A x
  Amg::Lev& C = *A.lev[k + 1];
  AmgLevelBufs<T>& CB = C.bufs<T>();
  const DevCsr<T> R{L.m, L.n, L.Rptr.p, L.Rcol.p, B.Rval.p, L.lanesR};
  amg::spmv<T>(R, B.r.p, CB.b.p, nullptr, 0, stream_);
  amg_vcycle<T>(k + 1);
  const DevCsr<T> P{L.n, L.m, L.Pptr.p, L.Pcol.p, B.Pval.p, L.lanesP};
  amg::spmv<T>(P, CB.x.p, B.x.p, nullptr, 1, stream_);  // x += P x_c
  smooth(false);
}

template <typename T>
void PoissonSolver::coarse_amg(int l) {
  Level& lv = *levels_[l];
  Work<T>& w = lv.work<T>();
  Amg& A = *amg_;
  const int64_t n = (int64_t)lv.L.Np * stride_;
  const int pre = l < (int)pmg_.pre_it.size() ? pmg_.pre_it[l] : 1;
  const int post = l < (int)pmg_.post_it.size() ? pmg_.post_it[l] : 1;
  if (pre <= 0) fill<T>(n, w.x.p, (T)0, stream_);
  for (int s = 0; s < pre; ++s) chebyshev<T>(l, w.b.p, w.x.p, s == 0);
  apply_level<T>(l, w.x.p, w.ax.p);
  ++smoother_applies_;
  axpby<T>(n, (T)1, w.b.p, (T)-1, w.ax.p, stream_);  // ax = b - A x
  Amg::Lev& V = *A.lev[0];
  AmgLevelBufs<T>& VB = V.bufs<T>();
  // b_v = P_v^T (b - A x); then AMG cycles on A_v x_v = b_v
  amg::vertex_restrict<T>(A.nv, A.vptr.p, A.vlist.p, w.ax.p, VB.b.p, stream_);
  amg_vcycle<T>(0);
  if (pmg_.amg_cycles > 1) {
    // further cycles on the vertex residual: x_v += V(b_v - A_v x_v)
    const DevCsr<T> Am{V.n, V.n, V.Aptr.p, V.Acol.p, VB.Aval.p, V.lanesA};
    copy<T>(V.n, VB.b.p, VB.b0.p, stream_);
    copy<T>(V.n, VB.x.p, VB.acc.p, stream_);
    for (int c = 1; c < pmg_.amg_cycles; ++c) {
      amg::spmv<T>(Am, VB.acc.p, VB.b.p, VB.b0.p, 2, stream_);
      amg_vcycle<T>(0);
      axpby<T>(V.n, (T)1, VB.x.p, (T)1, VB.acc.p, stream_);
    }
    copy<T>(V.n, VB.acc.p, VB.x.p, stream_);
  }
  amg::vertex_prolong<T>(lv.L.Np, mesh_.n_cells, stride_, A.cv.p, VB.x.p, w.x.p, true, stream_);
  for (int s = 0; s < post; ++s) chebyshev<T>(l, w.b.p, w.x.p, false);
}

template <typename T>
void PoissonSolver::vcycle(int l) {
  Level& lv = *levels_[l];
  Work<T>& w = lv.work<T>();
  if (l + 1 == (int)levels_.size()) {
    if (amg_) coarse_amg<T>(l);
    else coarse_cg<T>(l);
    return;
  }
  const int64_t n = (int64_t)lv.L.Np * stride_;
  const int pre = l < (int)pmg_.pre_it.size() ? pmg_.pre_it[l] : 1;
  const int post = l < (int)pmg_.post_it.size() ? pmg_.post_it[l] : 1;
  if (pre <= 0) fill<T>(n, w.x.p, (T)0, stream_);
  for (int s = 0; s < pre; ++s) chebyshev<T>(l, w.b.p, w.x.p, s == 0);
  apply_level<T>(l, w.x.p, w.ax.p);
  ++smoother_applies_;
  axpby<T>(n, (T)1, w.b.p, (T)-1, w.ax.p, stream_);  // ax = b - A x
  Level& cl = *levels_[l + 1];
  Work<T>& cw = cl.work<T>();
  launch_elem_matvec<T>(mesh_.dim, cl.N, lv.N, mesh_.n_cells, stride_, w.Pt.p, w.ax.p, cw.b.p,
                        false, stream_);
  vcycle<T>(l + 1);
  launch_elem_matvec<T>(mesh_.dim, lv.N, cl.N, mesh_.n_cells, stride_, w.P.p, cw.x.p, w.x.p, true,
                        stream_);
  for (int s = 0; s < post; ++s) chebyshev<T>(l, w.b.p, w.x.p, false);
}

void PoissonSolver::precondition(const double* r, double* z) {
  const int64_t n = (int64_t)np() * stride_;
  Level& l0 = *levels_[0];
  if (type_ == SolverType::JACOBI_PCG) {
    jacobi<double>(n, r, l0.w64.inv.p, z, stream_);
  } else if (type_ == SolverType::BLOCK_JACOBI_PCG) {
    block_apply(np(), mesh_.n_cells, stride_, blk_.p, r, z, stream_);
  } else if (type_ == SolverType::INV_MASS_PCG) {
    launch_inv_mass(np(), mesh_.n_cells, stride_, J_.p, minv_.p, inv_mass_factor_,
                    inv_mass_recp_.empty() ? nullptr : inv_mass_recp_.p, r, z, stream_);
  } else if (prec_ == Precision::FP64) {
    // the flexible CG keeps its residual in the finest level's right-hand side (which the
    // V-cycle only reads) and takes the V-cycle's result by trading buffers with z (every
    // V-cycle writes its iterate before reading it)
    if (r != l0.w64.b.p) copy<double>(n, r, l0.w64.b.p, stream_);
    vcycle<double>(0);
    if (z == z_.p) {
      std::swap(z_.p, l0.w64.x.p);
      z = z_.p;
    } else {
      copy<double>(n, l0.w64.x.p, z, stream_);
    }
  } else {
    convert<double, float>(n, r, l0.w32.b.p, stream_);
    vcycle<float>(0);
    convert<float, double>(n, l0.w32.x.p, z, stream_);
  }
  if (nullspace_) {
    RedScratch rs{partial_.p};
    vsum<double>(n, z, scal_.p, S_MEAN, rs, stream_);
    sub_mean<double>(np(), mesh_.n_cells, stride_, z, scal_.p, S_MEAN,
                     (double)np() * mesh_.n_cells, stream_);
  }
}

bool PoissonSolver::mixed_refinement() const {
  return prec_ == Precision::MIXED && type_ == SolverType::PMULTIGRID;
}

// Flexible CG with the FP32 V-cycle, FP32 operator and FP32 Krylov vectors; the iterate x,
// the right-hand side and every true residual r = b - A x are FP64, and the inner products and
// step lengths are accumulated and formed in FP64 on the device. The recursively updated FP32
// residual is replaced by the true FP64 residual (iterative refinement without a restart of
// the Krylov recurrence) whenever it has fallen by kRefineRatio since the last replacement or
// below the tolerance; the solve stops only on a true residual.
int PoissonSolver::fcg_mixed(double tol, int* iters, double* rr_out, int64_t* fine_applies) {
  constexpr double kRefineRatio = 1e-2;
  const int64_t n = (int64_t)np() * stride_;
  RedScratch rs{partial_.p};
  Level& l0 = *levels_[0];
  Work<float>& w = l0.w32;
  float* r = w.b.p;  // the V-cycle reads its right-hand side from the level's b
  auto true_residual = [&]() {
    apply_level<double>(0, x_.p, Ap_.p);
    ++*fine_applies;
    residual_f32(n, b_.p, Ap_.p, r, scal_.p, S_RR, rs, stream_);
    return read_scalar(S_RR);
  };
  auto precondition32 = [&]() {
    vcycle<float>(0);
    std::swap(z32_, w.x);  // the V-cycle's result becomes z; w.x is rewritten by the next one
    if (nullspace_) {
      vsum<float>(n, z32_.p, scal_.p, S_MEAN, rs, stream_);
      sub_mean<float>(np(), mesh_.n_cells, stride_, z32_.p, scal_.p, S_MEAN,
                      (double)np() * mesh_.n_cells, stream_);
    }
  };
  int it = 0;
  double rr = true_residual();
  *rr_out = rr;
  if (!std::isfinite(rr)) return -3;
  if (std::sqrt(rr) <= tol) return 0;
  double last_true = std::sqrt(rr);
  precondition32();
  dot<float>(n, r, z32_.p, nullptr, nullptr, 1, scal_.p, S_RZ, rs, stream_);
  copy<float>(n, z32_.p, p32_.p, stream_);
  int status = 1;
  while (it < max_iter_) {
    if (matrix_pass(0) && !oi_) {
      // the matrix-core element pass has no fused dot
      apply_level<float>(0, p32_.p, Ap32_.p);
      dot<float>(n, p32_.p, Ap32_.p, nullptr, nullptr, 1, scal_.p, S_PAP, rs, stream_);
    } else {
      apply_level<float>(0, p32_.p, Ap32_.p, partial_.p);  // p.Ap fused into the element pass
      finalize(apply_blocks(mesh_.n_cells), 1, partial_.p, scal_.p, S_PAP, stream_);
    }
    cg_update_xr_mixed(n, x_.p, r, p32_.p, Ap32_.p, scal_.p, S_RZ, S_PAP, S_RR, rs, stream_);
    ++it;
    read_scalar(0);
    const double pap = host_scal_[S_PAP];
    rr = host_scal_[S_RR];
    if (!(pap > 0) || !std::isfinite(rr) || !std::isfinite(pap)) { status = -3; break; }
    if (std::sqrt(rr) <= tol || std::sqrt(rr) <= kRefineRatio * last_true) {
      rr = true_residual();
      *rr_out = rr;
      if (!std::isfinite(rr)) { status = -3; break; }
      if (std::sqrt(rr) <= tol) { status = 0; break; }
      last_true = std::sqrt(rr);
    }
    if (it >= max_iter_) break;
    std::swap(z32_, zold32_);
    precondition32();
    dot_rz_flex<float>(n, r, z32_.p, zold32_.p, scal_.p, S_RZN, rs, stream_);  // and S_BNUM
    cg_update_p<float>(n, p32_.p, z32_.p, scal_.p, S_BNUM, S_RZ, stream_);
    scal_copy(scal_.p, S_RZ, S_RZN, stream_);
  }
  if (status == 1) {
    // not converged within max_iter: report the true residual of the final iterate
    rr = true_residual();
    *rr_out = rr;
    if (!std::isfinite(rr)) status = -3;
    else if (std::sqrt(rr) <= tol) status = 0;
  }
  *iters = it;
  return status;
}

void PoissonSolver::import_vec(const double* src, double* dst, Layout layout, int rows) {
  const int64_t E = mesh_.n_cells;
  if (!src) throw InvalidArgument("null vector");
  if (layout == Layout::HOST_AOS) {
    DBuf<double> tmp;
    tmp.alloc((size_t)rows * E, false);
    INSP_HIP(hipMemcpyAsync(tmp.p, src, (size_t)rows * E * sizeof(double), hipMemcpyHostToDevice,
                            stream_));
    soa_from_aos(rows, E, stride_, tmp.p, dst, stream_);
    INSP_HIP(hipStreamSynchronize(stream_));
  } else if (layout == Layout::DEVICE_AOS) {
    soa_from_aos(rows, E, stride_, src, dst, stream_);
  } else if (dat_stride() == stride_) {
    if (src != dst)
      INSP_HIP(hipMemcpyAsync(dst, src, (size_t)rows * stride_ * sizeof(double),
                              hipMemcpyDeviceToDevice, stream_));
    zero_padding<double>(rows, E, stride_, dst, stream_);
  } else {
    // the caller's dat stride (OP2 pads its SoA sets by its own rule)
    INSP_HIP(hipMemcpy2DAsync(dst, stride_ * sizeof(double), src, dat_stride() * sizeof(double),
                              E * sizeof(double), rows, hipMemcpyDeviceToDevice, stream_));
    zero_padding<double>(rows, E, stride_, dst, stream_);
  }
}

void PoissonSolver::export_vec(const double* src, double* dst, Layout layout, int rows) {
  const int64_t E = mesh_.n_cells;
  if (!dst) throw InvalidArgument("null vector");
  if (layout == Layout::HOST_AOS) {
    DBuf<double> tmp;
    tmp.alloc((size_t)rows * E, false);
    aos_from_soa(rows, E, stride_, src, tmp.p, stream_);
    INSP_HIP(hipMemcpyAsync(dst, tmp.p, (size_t)rows * E * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    INSP_HIP(hipStreamSynchronize(stream_));
  } else if (layout == Layout::DEVICE_AOS) {
    aos_from_soa(rows, E, stride_, src, dst, stream_);
  } else if (dat_stride() != stride_) {
    // only the caller's cells are written; its padding entries are left as they are
    INSP_HIP(hipMemcpy2DAsync(dst, dat_stride() * sizeof(double), src, stride_ * sizeof(double),
                              E * sizeof(double), rows, hipMemcpyDeviceToDevice, stream_));
  } else if (src != dst) {
    INSP_HIP(hipMemcpyAsync(dst, src, (size_t)rows * stride_ * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
  }
}

void PoissonSolver::set_dat_stride(int64_t s) {
  if (s != 0 && s < mesh_.n_cells)
    throw InvalidArgument("dat stride must be 0 (the solver's own) or at least the number of cells");
  dat_stride_ = s;
}

int PoissonSolver::solve(const double* rhs, double* ans, Layout layout, SolveStats* stats) {
  if (!initialised_) init();
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = (int64_t)np() * stride_;
  RedScratch rs{partial_.p};
  smoother_applies_ = 0;
  int64_t fine_applies = 0;
  import_vec(rhs, b_.p, layout, np());
  if (have_bc_)
    launch_apply_bc(mesh_view(), level_view(0, false), bc_.p, var_factor_ ? levels_[0]->w64.k.p : nullptr,
                    b_.p, stream_);
  if (nullspace_) {
    vsum<double>(n, b_.p, scal_.p, S_MEAN, rs, stream_);
    sub_mean<double>(np(), mesh_.n_cells, stride_, b_.p, scal_.p, S_MEAN,
                     (double)np() * mesh_.n_cells, stream_);
  }
  import_vec(ans, x_.p, layout, np());
  dot<double>(n, b_.p, b_.p, nullptr, nullptr, 1, scal_.p, S_BB, rs, stream_);
  const double bnorm = std::sqrt(read_scalar(S_BB));
  SolveStats st;
  st.bnorm = bnorm;
  int status = 1;
  int it = 0;
  double rr_true = 0;
  if (!std::isfinite(bnorm)) {
    status = -3;
  } else if (mixed_refinement()) {
    status = fcg_mixed(std::max(rtol_ * bnorm, atol_), &it, &rr_true, &fine_applies);
  } else {
    const double tol = std::max(rtol_ * bnorm, atol_);
    // FP64 p-multigrid: the residual lives in the finest level's right-hand side (no copy into
    // the V-cycle)
    double* const r = type_ == SolverType::PMULTIGRID ? levels_[0]->w64.b.p : r_.p;
    for (int restart = 0;; ++restart) {
      apply_level<double>(0, x_.p, Ap_.p);
      ++fine_applies;
      residual(n, b_.p, Ap_.p, r, scal_.p, S_RR, rs, stream_);
      rr_true = read_scalar(S_RR);
      if (!std::isfinite(rr_true)) { status = -3; break; }
      if (std::sqrt(rr_true) <= tol) { status = 0; break; }
      if (it >= max_iter_ || restart > 50) { status = 1; break; }
      precondition(r, z_.p);
      dot<double>(n, r, z_.p, nullptr, nullptr, 1, scal_.p, S_RZ, rs, stream_);
      copy<double>(n, z_.p, p_.p, stream_);
      bool breakdown = false;
      while (it < max_iter_) {
        apply_level<double>(0, p_.p, Ap_.p, partial_.p);  // p.Ap fused into the element pass
        ++fine_applies;
        finalize(apply_blocks(mesh_.n_cells), 1, partial_.p, scal_.p, S_PAP, stream_);
        cg_update_xr<double>(n, x_.p, r, p_.p, Ap_.p, nullptr, nullptr, scal_.p, S_RZ, S_PAP,
                             S_RR, rs, stream_);
        ++it;
        read_scalar(0);
        const double pap = host_scal_[S_PAP], rr = host_scal_[S_RR];
        if (!(pap > 0) || !std::isfinite(rr) || !std::isfinite(pap)) { breakdown = true; break; }
        if (std::sqrt(rr) <= tol) break;
        std::swap(z_.p, zold_.p);
        precondition(r, z_.p);
        dot_rz_flex<double>(n, r, z_.p, zold_.p, scal_.p, S_RZN, rs, stream_);  // and S_BNUM
        cg_update_p<double>(n, p_.p, z_.p, scal_.p, S_BNUM, S_RZ, stream_);
        scal_copy(scal_.p, S_RZ, S_RZN, stream_);
      }
      if (breakdown) { status = -3; break; }
    }
  }
  if (status != -3 && nullspace_) {
    vsum<double>(n, x_.p, scal_.p, S_MEAN, rs, stream_);
    sub_mean<double>(np(), mesh_.n_cells, stride_, x_.p, scal_.p, S_MEAN,
                     (double)np() * mesh_.n_cells, stream_);
  }
  export_vec(x_.p, ans, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
  const auto t1 = std::chrono::steady_clock::now();
  st.iterations = it;
  st.converged = status == 0;
  st.abs_residual = std::sqrt(rr_true);
  st.rel_residual = bnorm > 0 ? st.abs_residual / bnorm : 0.0;
  st.seconds = std::chrono::duration<double>(t1 - t0).count();
  st.fine_applies = fine_applies;
  st.smoother_applies = smoother_applies_;
  if (stats) *stats = st;
  return status;
}

void PoissonSolver::mult(const double* in, double* out, Layout layout, bool fp32) {
  if (!initialised_) init();
  const int64_t n = (int64_t)np() * stride_;
  import_vec(in, b_.p, layout, np());
  if (!fp32) {
    apply_level<double>(0, b_.p, Ap_.p);
  } else {
    Level& l0 = *levels_[0];
    DBuf<float> xf, yf;
    xf.alloc(n);
    yf.alloc(n);
    if (var_factor_ && l0.w32.k.empty()) {
      l0.w32.k.alloc(n);
      convert<double, float>(n, l0.w64.k.p, l0.w32.k.p, stream_);
    }
    convert<double, float>(n, b_.p, xf.p, stream_);
    apply_level<float>(0, xf.p, yf.p);
    convert<float, double>(n, yf.p, Ap_.p, stream_);
    INSP_HIP(hipStreamSynchronize(stream_));
  }
  export_vec(Ap_.p, out, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::mult_noalloc(const double* in, double* out, Layout layout) {
  if (!initialised_)
    throw InvalidArgument("the allocation-free apply needs an initialised solver (ins_poisson_init)");
  if (layout == Layout::HOST_AOS) throw InvalidArgument("the allocation-free apply takes device vectors");
  if (!in || !out) throw InvalidArgument("null vector");
  const Level& l0 = *levels_[0];
  const size_t n_tr = (size_t)l0.L.Nf * l0.L.Nfp * stride_;  // FP64 face traces
  const size_t n_vec = (size_t)np() * stride_;
  // the solver's own SoA layout runs on the caller's vectors; any other is staged in the scratch
  const bool direct = layout == Layout::DEVICE_SOA && dat_stride() == stride_ && in != out;
  const size_t need = (n_tr + (direct ? 0 : 2 * n_vec)) * sizeof(double);
  if (need > kNoAllocScratchBytes)
    throw InvalidArgument("mesh too large for the allocation-free apply: needs " + std::to_string(need) +
                          " bytes of scratch, the limit is " + std::to_string(kNoAllocScratchBytes));
  double* scr = noalloc_scratch();
  if (direct) {
    apply_level<double>(0, in, out, nullptr, scr);
  } else {
    double* u = scr + n_tr;
    double* y = u + n_vec;
    import_vec(in, u, layout, np());
    apply_level<double>(0, u, y, nullptr, scr);
    export_vec(y, out, layout, np());
  }
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::diag(double* out, Layout layout) {
  if (!initialised_) init();
  export_vec(levels_[0]->diag.p, out, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::apply_bc(double* rhs, Layout layout) {
  if (!initialised_) init();
  import_vec(rhs, b_.p, layout, np());
  if (have_bc_)
    launch_apply_bc(mesh_view(), level_view(0, false), bc_.p,
                    var_factor_ ? levels_[0]->w64.k.p : nullptr, b_.p, stream_);
  export_vec(b_.p, rhs, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::mass(const double* in, double* out, Layout layout) {
  if (!initialised_) init();
  import_vec(in, b_.p, layout, np());
  launch_mass(mesh_view(), level_view(0, false), b_.p, Ap_.p, stream_);
  export_vec(Ap_.p, out, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::pressure_rhs(double b0, double b1, double dt, const double* dPdN, double* dPdNOld,
                                 double* div, Layout layout) {
  if (!(dt > 0) || !std::isfinite(b0) || !std::isfinite(b1)) throw InvalidArgument("invalid rhs coefficients");
  if (!dPdN || !dPdNOld || !div) throw InvalidArgument("null vector");
  const int Np = np(), NF = (mesh_.dim + 1) * nfp();
  if (Mfine_.empty()) {
    Mfine_.upload(fine_re_.M.data(), fine_re_.M.size());
    Efine_.upload(fine_re_.EMAT.data(), fine_re_.EMAT.size());
  }
  if (layout == Layout::DEVICE_SOA && dat_stride() == stride_) {
    launch_pressure_rhs(mesh_.dim, N_, mesh_.n_cells, stride_, J_.p, Mfine_.p, Efine_.p, b0, b1, dt, dPdN,
                        dPdNOld, div, stream_);
    INSP_HIP(hipStreamSynchronize(stream_));
    return;
  }
  DBuf<double> g0, g1, v;
  g0.alloc((size_t)NF * stride_);
  g1.alloc((size_t)NF * stride_);
  v.alloc((size_t)Np * stride_);
  import_vec(dPdN, g0.p, layout, NF);
  import_vec(dPdNOld, g1.p, layout, NF);
  import_vec(div, v.p, layout, Np);
  launch_pressure_rhs(mesh_.dim, N_, mesh_.n_cells, stride_, J_.p, Mfine_.p, Efine_.p, b0, b1, dt, g0.p,
                      g1.p, v.p, stream_);
  export_vec(g1.p, dPdNOld, layout, NF);
  export_vec(v.p, div, layout, Np);
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::set_history_size(int n) {
  if (n < 1 || n > 8) throw InvalidArgument("history size must be in 1..8");
  hist_max_ = n;
  while ((int)hist_.size() > hist_max_) hist_.erase(hist_.begin());
}

void PoissonSolver::history_push(double t, const double* x, Layout layout) {
  if (!std::isfinite(t)) throw InvalidArgument("invalid time");
  for (const Snapshot& h : hist_)
    if (h.t == t) throw InvalidArgument("history already holds this time");
  Snapshot s;
  s.t = t;
  s.x.alloc((size_t)np() * stride_);
  import_vec(x, s.x.p, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
  hist_.push_back(std::move(s));
  while ((int)hist_.size() > hist_max_) hist_.erase(hist_.begin());
}

void PoissonSolver::extrapolate(double t, double* ans, Layout layout) {
  if (hist_.empty()) return;  // keep the caller's guess
  const int n = (int)hist_.size();
  const int64_t len = (int64_t)np() * stride_;
  DBuf<double> acc;
  acc.alloc((size_t)len);
  for (int i = 0; i < n; ++i) {
    double w = 1.0;  // Lagrange basis polynomial of node i at t
    for (int j = 0; j < n; ++j)
      if (j != i) w *= (t - hist_[j].t) / (hist_[i].t - hist_[j].t);
    axpby<double>(len, w, hist_[i].x.p, 1.0, acc.p, stream_);
  }
  export_vec(acc.p, ans, layout, np());
  INSP_HIP(hipStreamSynchronize(stream_));
}

void PoissonSolver::node_coords(double* xyz) const { mesh_node_coords(mesh_, fine_re_, xyz); }

const HostCsr& PoissonSolver::vertex_matrix() const {
  if (!amg_) throw InvalidArgument("no vertex coarse space: configure the AMG coarse solve and init");
  return amg_->Av;
}

std::vector<int> PoissonSolver::amg_level_sizes() const {
  std::vector<int> s;
  if (amg_)
    for (auto& L : amg_->lev) s.push_back(L->n);
  return s;
}

double PoissonSolver::level_lambda_max(int l) const {
  if (l < 0 || l >= (int)levels_.size()) throw InvalidArgument("level out of range");
  return levels_[l]->lmax;
}

}  // namespace insp
