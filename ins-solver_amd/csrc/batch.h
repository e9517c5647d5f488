// Batched Poisson solves: many independent problems of uneven size and order
// advanced in lockstep by one set of launches.
//
// Every problem owns one device arena (mesh, factors, Krylov state stay
// resident). Its cells are cut into 256-cell tiles that never straddle
// problems; a device tile table (problem slot, first cell) drives every batched
// kernel, which picks the problem's order at run time. Reductions are written
// per tile and added per problem in tile order, so a problem's arithmetic is
// bitwise the same alone or in any mix. The outer solver is the reference's
// flexible PCG (viscous_solver.cpp:158-266) in FP64; the preconditioner is a
// two-level p-multigrid V-cycle (order N -> 1, the reference's "orders = N,1",
// examples/2d/vortex/config.ini) with Chebyshev smoothing and Jacobi-CG on
// order 1, in FP64 or FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <map>
#include <memory>
#include <vector>

#include "mesh.h"

namespace insp {

using HostMeshPtr = std::shared_ptr<HostMesh>;

constexpr int kTile = 256;  // stride granularity; cell tiles are 64, 128 or 256 cells

// Vector fields of a problem, indexed by id in the descriptor.
enum BatchVec : int {
  BV_B = 0, BV_X, BV_R, BV_Z, BV_P, BV_AP, BV_TRD,   // FP64 outer flexible CG (+ traces)
  BV_VX, BV_VB, BV_VAX, BV_VD, BV_TRT,               // V-cycle, order N, precision T
  BV_CX, BV_CB, BV_CR, BV_CZ, BV_CP, BV_CAP,         // V-cycle, order 1, precision T
  BV_COUNT
};

// Per-problem scalar slots.
enum BatchScal : int {
  BS_RZ = 0, BS_PAP, BS_RR, BS_RZO, BS_RZN, BS_BB, BS_MEAN,
  BS_CA, BS_CPAP, BS_CB,
  BS_TOP, BS_BOT, BS_NRM,
  // device-side convergence mask: a problem whose state is not RUNNING is frozen (every
  // batched kernel skips it) until the host restarts or retires it
  BS_STATE, BS_ITER, BS_TOL, BS_MAXIT,
  BS_COUNT = 20
};
enum BatchState : int { BST_RUNNING = 0, BST_TOL = 1, BST_BREAKDOWN = 2, BST_MAXIT = 3 };

constexpr int kAmgLevels = 8;   // deepest AMG hierarchy of one problem in a batch
constexpr int kAmgMaxCoarse = 256;

// One level of a problem's vertex-space AMG hierarchy (values in the engine's
// V-cycle precision).
struct BatchAmgLevel {
  int n, m;                      // rows, rows of the next level (0 on the last)
  const int *Aptr, *Acol, *Pptr, *Pcol, *Rptr, *Rcol;
  const void *Aval, *Pval, *Rval, *inv;
  void *x, *b, *r, *d;
  double ch1[4], ch2[4];         // Chebyshev step coefficients
};

struct BatchAmg {
  int nlev, nv, nc, pad;         // levels, vertices, size of the dense last level
  const int* cv;                 // order-1 SoA entry -> vertex
  const int* vptr;               // vertex -> SoA entries (gather list)
  const int* vlist;
  const void* cinv;              // dense symmetric inverse on the last level
  double oc1[8], oc2[8];         // Chebyshev coefficients of the order-1 DG level
  BatchAmgLevel lev[kAmgLevels];
};

struct BatchProb {
  int N, Np, Np1, var, nullspace, pad;
  int64_t E, S;                  // cells, stride (multiple of kTile)
  double dofs;                   // E * Np (mean projection)
  const double* Kt;
  const double* J;
  const int* nbr;
  const int* code;
  const int* bctype;
  const double* tauF;            // penalty * sJ, order N
  const double* tauC;            // penalty * sJ, order 1 (penalty order N)
  const void* kF[2];             // factor at order N / 1, [0] double [1] float; null = 1
  const void* kC[2];
  const void* invF[2];           // inverse diagonals
  const void* invC[2];
  void* v[BV_COUNT];
  double* scal;                  // BS_COUNT doubles
  double c1[8], c2[8];           // Chebyshev step coefficients (order N level)
  BatchAmg amg;                  // vertex-space coarse solve (nlev = 0: Jacobi-CG on order 1)
};

struct BatchTables {
  const double* C64[5];
  const float* C32[5];
  const int* tab[5];
  const double* P64[5];          // prolongation order 1 -> N (Np x Np1)
  const float* P32[5];
  const double* Pt64[5];         // restriction (Np1 x Np)
  const float* Pt32[5];
};

struct BatchView {
  const BatchProb* P;            // descriptors by slot
  const int2* tiles;             // (slot, first cell)
  const int* ntiles;             // device count (<= grid)
  const int* probs;              // active slots
  const int2* ranges;            // tile range per slot
  const int* nprobs;             // device count
  int tile_cap, prob_cap;        // grid sizes
  double* partial;               // 4 * tile_cap doubles
  int dim;
  int tile;                      // cells per tile = threads per block (64, 128, 256)
  // Tiles are grouped by problem order: order o owns tiles [ooff[o], ooff[o] + ocount[o]) of
  // the table (device values), launched with a grid of ocap[o] blocks.
  const int* ocount;
  const int* ooff;
  int ocap[5];
  // AMG row tiles: level k lists 256-row tiles (slot, first row) of every problem that
  // has level k, at ltiles[k * ltile_cap + t], with lcount[k] entries.
  const int2* ltiles;
  const int* lcount;
  int ltile_cap;
};

namespace batch {
// level: 0 = order N, 1 = order 1. Field ids select BatchVec members.
template <typename T>
void trace(const BatchView& b, const BatchTables& t, int in, int tr, hipStream_t s);
template <typename T>
void apply(const BatchView& b, const BatchTables& t, int level, int in, int tr, int out, bool dot,
           hipStream_t s);
void finalize(const BatchView& b, int nvals, int slot, hipStream_t s);
// After an iteration's |r|^2: count the iteration and freeze problems that reached their
// tolerance or iteration limit or broke down (p.Ap <= 0, non-finite values).
void mask(const BatchView& b, hipStream_t s);
// rec holds n records (slot, state, iterations, tolerance, iteration limit): writes the state,
// and with `limits` the other three, into each slot's scalars.
void set_state(double* scal, int n, const double* rec, bool limits, hipStream_t s);
// value k of the per-tile partials: sum of a_k*b_k over the tile (level rows).
template <typename T>
void dot(const BatchView& b, int level, int a, int bb, int c, int d, int nvals, hipStream_t s);
void fcg_update_xr(const BatchView& b, hipStream_t s);     // partials: |r|^2, r.z(old)
void fcg_update_p(const BatchView& b, hipStream_t s);
void scal_copy(const BatchView& b, int dst, int src, hipStream_t s);
void residual(const BatchView& b, hipStream_t s);           // r = b - Ap, partial |r|^2
void to_t(const BatchView& b, bool fp32, int src_d, int dst, hipStream_t s);   // double -> T
void from_t(const BatchView& b, bool fp32, int src, int dst_d, hipStream_t s); // T -> double
// level 0: order-N fields (VB, VAX, VD, VX, invF, c1/c2); level 1: order-1 fields
// (CB, CAP, CZ, CX, invC, amg.oc1/oc2).
template <typename T>
void cheb(const BatchView& b, int level, int step, bool zero_init, hipStream_t s);
template <typename T>
void sub_from_b(const BatchView& b, int level, hipStream_t s);  // ax = b - ax on that level
// Vertex-space AMG (problems with amg.nlev > 0).
template <typename T>
void vertex_restrict(const BatchView& b, hipStream_t s);   // lev[0].b = P_v^T CAP
template <typename T>
void vertex_prolong(const BatchView& b, hipStream_t s);    // CX += P_v lev[0].x
template <typename T>
void amg_spmv(const BatchView& b, int k, bool residual, hipStream_t s);  // r = A x or b - A x
template <typename T>
void amg_cheb(const BatchView& b, int k, int step, bool zero_init, hipStream_t s);
template <typename T>
void amg_restrict(const BatchView& b, int k, hipStream_t s);  // lev[k+1].b = R lev[k].r
template <typename T>
void amg_prolong(const BatchView& b, int k, hipStream_t s);   // lev[k].x += P lev[k+1].x
template <typename T>
void amg_dense(const BatchView& b, int k, hipStream_t s);     // last level k: x = Cinv b

template <typename T>
void restrict_(const BatchView& b, const BatchTables& t, hipStream_t s);
template <typename T>
void prolong(const BatchView& b, const BatchTables& t, hipStream_t s);
template <typename T>
void cg_init(const BatchView& b, hipStream_t s);            // partial cr.cz
template <typename T>
void cg_xr(const BatchView& b, int num, hipStream_t s);     // partial cr.cz
template <typename T>
void cg_p(const BatchView& b, int num, int den, hipStream_t s);
void sum_d(const BatchView& b, int field, hipStream_t s);   // partial sum (double field)
void sub_mean_d(const BatchView& b, int field, hipStream_t s);
void rand_d(const BatchView& b, int level, int field, hipStream_t s);
void inv_apply_d(const BatchView& b, int level, int field, hipStream_t s);  // field *= inv (double)
void scale_d(const BatchView& b, int field, int slot, hipStream_t s); // field /= sqrt(scal)
}  // namespace batch

struct BatchConfig {
  int dim = 2;
  bool fp32 = true;          // precision of the V-cycle (the outer flexible CG is FP64)
  int cheb_order = 3;        // Chebyshev degree of one smoothing application
  int pre_it = 1, post_it = 1;
  double eig_safety = 1.1, cheb_min_factor = 0.1;
  int max_iter_eig = 15;
  int coarse_iters = 40;     // Jacobi-CG iterations on order 1
  bool amg = true;           // vertex-space AMG below order 1 instead of Jacobi-CG
  int amg_cheb = 2;          // Chebyshev degree on the AMG levels
  bool use_graph = true;     // replay captured iterations
  int chunk = 8;             // lockstep iterations per replay between host reads
  int tile = 0;              // cells per tile: 64, 128, 256, or 0 = autotuned
};

struct ProblemSpec {
  HostMeshPtr mesh;          // owned
  int order = 2;
  std::vector<int> bc_types; // per bface; empty = all Dirichlet
  std::vector<double> factor;// cells x Np (AoS); empty = constant 1
  bool nullspace = false;
  double rtol = 1e-8, atol = 0.0;
  int max_iter = 500;
  std::vector<double> rhs;   // cells x Np (AoS), mass-weighted
  bool rhs_nodal = false;    // rhs holds nodal f instead: b = M f is formed on the device
};

enum class ProbStatus { RUNNING = 0, CONVERGED = 1, NOT_CONVERGED = 2, BREAKDOWN = 3 };

struct ProbResult {
  ProbStatus status = ProbStatus::RUNNING;
  int iterations = 0;
  double rel_residual = 0, bnorm = 0;
  double seconds = 0;        // add -> finished, wall clock
  int64_t fine_applies = 0;  // FP64 applications of the order-N operator
  int64_t vcycles = 0;       // preconditioner applications
  std::vector<double> x;     // cells x Np (AoS), filled when finished
};

class BatchEngine {
 public:
  explicit BatchEngine(const BatchConfig& cfg, hipStream_t stream = nullptr);
  ~BatchEngine();
  BatchEngine(const BatchEngine&) = delete;
  BatchEngine& operator=(const BatchEngine&) = delete;

  // Device bytes one problem occupies (admission by HBM budget).
  static size_t arena_bytes(int dim, int64_t cells, int order, bool var, bool fp32);
  // Upper estimate of the extra device bytes of the AMG coarse solve of one problem.
  static size_t amg_bytes_estimate(int dim, int64_t n_nodes, int64_t cells, bool fp32);
  // Uploads, sets up (penalty, diagonals, Chebyshev bounds) and starts a problem.
  // Returns its id. Set-up of several problems can be batched with add_many.
  std::vector<int> add_many(std::vector<ProblemSpec>&& specs);
  // k lockstep flexible-CG iterations of every running problem without a host round trip
  // (problems that converge on the way are frozen on the device); returns k.
  int iterate(int k = 1);
  // iterate in two halves, so that the work of several engines overlaps on the device:
  // iterate_launch enqueues k iterations and the read-back of the problems' state on the
  // engine's stream and returns at once (the number of iterations launched, 0 when no problem
  // runs); iterate_collect waits for them and retires the problems that finished.
  int iterate_launch(int k = 1);
  void iterate_collect();
  int running() const;
  int free_slots() const { return (int)free_slots_.size(); }
  const BatchConfig& config() const { return cfg_; }
  // Ids finished since the last call (their results are ready).
  std::vector<int> take_finished();
  const ProbResult& result(int id) const;
  void release(int id);      // frees the problem's device memory
  size_t bytes_in_use() const { return bytes_in_use_; }
  int64_t iterations_run() const { return iters_run_; }
  int tile() const { return tile_; }
  // Wall seconds of problem set-up so far: upload + penalties/diagonals, Chebyshev bounds,
  // AMG (probing + host hierarchy), start; and the host-hierarchy share of the AMG part.
  const double* setup_seconds() const { return setup_s_; }
  // Times the batched operator for each tile size on these problems (already added) and
  // returns the fastest; the choice is cached per (dim, precision) for the process.
  struct TuneResult {
    int tile;
    double ms[3];              // per tile size 64, 128, 256
  };
  static TuneResult last_tune(int dim, bool fp32);

 private:
  struct Prob;
  struct View;
  void build_tables();
  void upload_desc(int slot);
  void set_view(View& v, const std::vector<int>& slots);
  void precondition(View& v);
  void run_iteration(View& v);
  void start(View& v, const std::vector<int>& slots, bool restart);
  void finish(int slot, ProbStatus st);
  struct AmgPrep;
  struct ProbeScratch;
  void setup_amg_all(const std::vector<int>& slots, const std::vector<HostMeshPtr>& meshes);
  void amg_prepare(int slot, const HostMesh& m, AmgPrep& A);  // host, thread-safe
  void amg_probe(int slot, AmgPrep& A, ProbeScratch& ps);     // device
  void amg_build(int slot, AmgPrep& A);                       // host, thread-safe
  void amg_upload(int slot, AmgPrep& A);                      // device
  static void chebyshev_coefficients(double hi, double min_factor, int n, double* c1, double* c2);
  void read_scal(const std::vector<int>& slots);
  template <typename T>
  void vcycle(View& v);

  BatchConfig cfg_;
  int tile_ = 256;
  double setup_s_[5] = {0, 0, 0, 0, 0};
  bool tuned_ = false;
  void autotune(const std::vector<int>& slots);
  hipStream_t stream_;
  bool own_stream_ = false;
  int launched_ = 0;  // iterations launched and not yet collected
  BatchTables tables_{};
  std::vector<void*> table_bufs_;
  std::vector<std::unique_ptr<Prob>> probs_;   // by slot
  std::vector<int> free_slots_;
  BatchProb* d_desc_ = nullptr;                // device descriptors (max slots)
  double* d_scal_ = nullptr;                   // max slots x BS_COUNT
  double* h_scal_ = nullptr;                   // pinned mirror
  int max_slots_;
  std::unique_ptr<View> run_, work_;
  std::vector<int> running_slots_, finished_;
  size_t bytes_in_use_ = 0;
  int64_t iters_run_ = 0;
  std::map<int, hipGraphExec_t> graphs_;        // captured k-iteration replays, by k
  int graph_tile_cap_ = -1, graph_prob_cap_ = -1;
  double* h_stage_ = nullptr;                  // pinned staging of per-problem scalar writes
  double* d_stage_ = nullptr;
  void write_state(const std::vector<int>& slots, bool with_limits);
  void drop_graphs();
};

}  // namespace insp
