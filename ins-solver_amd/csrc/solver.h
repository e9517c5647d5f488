// C++ host side of the Poisson solve: the matrix-free SIPG operator and the
// linear solvers behind the C ABI. It mirrors the OP2-DG-Toolkit interface the
// reference calls (src/solvers/2d/ins_solver.cpp:91-96,205-207,324-325):
// PoissonMatrix::{set_bc_types, set_factor, mult} and
// LinearSolver::{set_bcs, set_nullspace, set_tol_and_iter, init, solve}.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "amg.h"
#include "dg_layout.h"
#include "kernels.h"
#include "mesh.h"
#include "oi.h"
#include "refelem.h"
#include "vec_kernels.h"

namespace insp {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct InvalidArgument : std::runtime_error {
  explicit InvalidArgument(const std::string& s) : std::runtime_error(s) {}
};

#define INSP_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t err_ = (call);                                                            \
    if (err_ != hipSuccess)                                                              \
      throw ::insp::HipError(std::string(#call) + ": " + hipGetErrorString(err_));       \
  } while (0)

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  // Allocates and zeroes. The zeroing runs on the null stream, which is not ordered with
  // the solver's non-blocking stream, so it is completed before returning.
  void alloc(size_t count, bool zero = true) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    INSP_HIP(hipMalloc(&p, count * sizeof(T)));
    if (zero) {
      INSP_HIP(hipMemset(p, 0, count * sizeof(T)));
      INSP_HIP(hipDeviceSynchronize());
    }
    n = count;
  }
  void upload(const T* h, size_t count) {
    alloc(count);
    INSP_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  bool empty() const { return p == nullptr; }
};

// Operator constants (D, B^T, face masses, W, M) and the index table
// (face masks, neighbour node matching, inverse masks) of one order, laid out by
// op_layout(); values in double.
void make_op_tables(const RefElement& re, std::vector<double>& C, std::vector<int>& tab);

enum class SolverType { JACOBI_PCG = 1, BLOCK_JACOBI_PCG = 2, INV_MASS_PCG = 3, PMULTIGRID = 4 };
// Solve on the coarsest p-multigrid order: Jacobi-CG there, or (AMG) Chebyshev
// smoothing there and a correction from the continuous vertex space solved by
// smoothed-aggregation AMG.
enum class CoarseType { JACOBI_CG = 0, AMG = 1 };
enum class Precision { FP64 = 0, MIXED = 1 };
// Element pass of the FP32 operator: AUTO = matrix cores at 3D order 4 (measured faster
// there), vector FMA elsewhere; VECTOR / MATRIX force one kernel where it exists.
enum class ElementPass { AUTO = 0, VECTOR = 1, MATRIX = 2 };
enum class Layout { HOST_AOS = 0, DEVICE_SOA = 1, DEVICE_AOS = 2 };

struct PMGConfig {
  std::vector<int> orders;        // finest first, e.g. {3,2,1}; empty = N..1
  std::vector<int> pre_it, post_it, cheb_orders;
  double eig_safety = 1.1;        // eigen_val_saftey_factor
  double cheb_min_factor = 0.1;   // lower Chebyshev bound = factor * upper
  int max_iter_eig = 15;          // max_iter_eig_approx
  int coarse_iters = 40;          // Jacobi-CG iterations on the coarsest order
  CoarseType coarse = CoarseType::JACOBI_CG;
  int amg_cycles = 1;             // AMG V-cycles per coarse correction
};

struct SolveStats {
  int iterations = 0;
  int converged = 0;
  double rel_residual = 0;        // true |b - A x| / |b| in FP64
  double abs_residual = 0;
  double bnorm = 0;
  double seconds = 0;
  int64_t fine_applies = 0;       // FP64 fine-level operator applications
  int64_t smoother_applies = 0;   // applications inside the preconditioner (all levels)
};

// Accumulated device time of the fine-level apply kernel, measured with HIP
// events on the solver's stream while profiling is enabled.
struct KernelTiming {
  double trace_ms = 0, apply_ms = 0;
  int64_t trace_launches = 0, apply_launches = 0;
};
// Device time of the operator applies on the coarsest p-multigrid level (below the
// finest), one launch in four timed, while profiling is enabled.
struct CoarseTiming {
  double apply_ms = 0;
  int64_t launches = 0;
};

class PoissonSolver {
 friend class VecViscousSolver;  // the vector (velocity) form shares the operator
 public:
  PoissonSolver(HostMesh mesh, int order, hipStream_t stream = nullptr);
  ~PoissonSolver();

  const HostMesh& mesh() const { return mesh_; }
  int order() const { return N_; }
  int np() const { return fine_re_.Np; }
  int nfp() const { return fine_re_.Nfp; }
  int64_t stride() const { return stride_; }
  // Stride of the caller's device SoA vectors (Layout::DEVICE_SOA), dat[i*stride + cell]:
  // 0 = the solver's own stride(); otherwise at least n_cells (an OP2 dat's padded set size).
  void set_dat_stride(int64_t s);
  int64_t dat_stride() const { return dat_stride_ ? dat_stride_ : stride_; }
  int64_t n_cells() const { return mesh_.n_cells; }
  hipStream_t stream() const { return stream_; }
  const RefElement& ref() const { return fine_re_; }

  void set_bc_types(const int* types, int64_t n);
  void set_factor(const double* k, Layout layout);  // nullptr = constant coefficient 1
  // Over-integration (FactorPoissonMatrixFreeDiagOI::mat_free_set_factor_oi /
  // mat_free_set_factor_surf_oi): the coefficient at the cubature points of the finest order's
  // rule (oi.h), cells x nv and cells x Nf*ns; on while both are set (null clears one). The
  // nodal factor of set_factor still sets the penalty, the boundary terms and the coarser
  // p-multigrid levels.
  void set_factor_oi(const double* kv, Layout layout);
  void set_factor_surf_oi(const double* ks, Layout layout);
  bool over_integrated() const { return oi_; }
  const OiRule& oi_rule();
  // Mass term of the viscous operator (FactorMMPoissonMatrixFree::set_mm_factor, per node;
  // MMPoissonMatrixFree::set_factor, constant). nullptr / 0 = none.
  void set_mm_factor(const double* mm, Layout layout);
  void set_mm_const(double mm);
  // Inverse-mass preconditioner scaling: times f (PETScInvMassSolver::setFactor,
  // ViscousSolver::set_inv_mass_factor) or divided by a per-node factor
  // (set_inv_mass_recp_factor; nullptr = use the scalar).
  void set_inv_mass_factor(double f) { inv_mass_factor_ = f; }
  void set_inv_mass_recp_factor(const double* f, Layout layout);
  // Pressure projection (DGMesh::grad_with_central_flux + ins_2d_proj_rhs): g[x] = grad_h p,
  // or g[x] = u[x] - dt grad_h p when u is given; vectors of the solver's order.
  void grad_central(const double* p, double* const g[3], const double* const u[3], double dt,
                    Layout layout);
  void set_bcs(const double* bc, Layout layout);    // bfaces x Nfp, nullptr = none
  void set_nullspace(bool ns) {
    if (ns != nullspace_ && pmg_.coarse == CoarseType::AMG) initialised_ = false;  // coarsest inverse
    nullspace_ = ns;
  }
  void set_tol_and_iter(double rtol, double atol, int max_iter);
  void set_solver(SolverType t) { type_ = t; initialised_ = false; }
  void set_precision(Precision p) { prec_ = p; initialised_ = false; }
  void set_element_pass(ElementPass p) { epass_ = p; initialised_ = false; }
  // whether level l's FP32 element pass runs on the matrix cores (after init)
  bool matrix_pass(int l) const;
  void set_pmg(const PMGConfig& c) { pmg_ = c; initialised_ = false; }
  const PMGConfig& pmg() const { return pmg_; }
  void init();

  // ans is the initial guess on entry (as in the reference) and the solution on exit.
  // Returns 0 converged, 1 not converged, -3 breakdown.
  int solve(const double* rhs, double* ans, Layout layout, SolveStats* stats);

  // out = A in on the finest order (FP64 or FP32 kernels).
  void mult(const double* in, double* out, Layout layout, bool fp32 = false);
  // FP64 apply that allocates nothing: face traces (and staging copies for other layouts) in a
  // static per-GPU scratch array; needs init() done and device vectors.
  void mult_noalloc(const double* in, double* out, Layout layout);
  void diag(double* out, Layout layout);
  // Inverses of the per-cell diagonal blocks (block-Jacobi solver type), host cells x Np x Np.
  void inv_block_diag(double* out);
  void apply_bc(double* rhs, Layout layout);
  void mass(const double* in, double* out, Layout layout);
  // The caller's pressure rhs (ins_pressure_rhs_2d.h + op2_gemv(LIFT) + mass,
  // src/solvers/2d/ins_solver.cpp:310-319): div <- J M (-div/dt + LIFT g),
  // g = b0 dPdN + b1 dPdNOld stored back into dPdNOld. Face data: cells x Nf*Nfp.
  void pressure_rhs(double b0, double b1, double dt, const double* dPdN, double* dPdNOld, double* div,
                    Layout layout);
  // Initial guess from earlier solutions (add_to_pr_history / initial_guess_extrapolation,
  // src/solvers/2d/ins_solver_base.cpp:1141-1155, src/solvers/3d/ins_solver.cpp:342-350):
  // Lagrange extrapolation in time through the stored (t, x) pairs.
  void set_history_size(int n);
  void history_push(double t, const double* x, Layout layout);
  int history_count() const { return (int)hist_.size(); }
  void extrapolate(double t, double* ans, Layout layout);
  void node_coords(double* xyz) const;  // host, dim x cells x Np (AoS per coordinate)
  double level_lambda_max(int l) const;
  int n_levels() const { return (int)levels_.size(); }
  // Vertex-space operator (CSR, host copy) and AMG level sizes; empty unless the AMG
  // coarse solve is configured and initialised.
  const HostCsr& vertex_matrix() const;
  std::vector<int> amg_level_sizes() const;

  void set_profiling(bool on) { profiling_ = on; }
  KernelTiming take_timing();
  CoarseTiming take_coarse_timing();

 private:
  struct Level;
  template <typename T>
  struct Work;

  void build_levels();
  void ensure_factor_levels();
  void compute_diagonals();
  void estimate_eigenvalues();
  template <typename T>
  void apply_level(int l, const T* u, T* out, double* partial = nullptr, T* tr = nullptr,
                   bool cells = false);
  // finest level: coefficient and mass-term factor (null when constant), face nodes per cell
  const double* fine_factor() const;
  const double* fine_mm() const;
  int fine_face_nodes() const;
  template <typename T>
  void vcycle(int l);
  template <typename T>
  void chebyshev(int l, const T* f, T* x, bool zero_init);
  template <typename T>
  void coarse_cg(int l);
  struct Amg;
  void build_coarse_space();
  void build_block_jacobi();
  template <typename T>
  void amg_vcycle(int k);
  template <typename T>
  void coarse_amg(int l);
  void precondition(const double* r, double* z);
  // FP32 inner iterations with FP64 iterate, true residuals and step lengths (p-multigrid
  // with Precision::MIXED); returns the solve status.
  bool mixed_refinement() const;
  int fcg_mixed(double tol, int* iters, double* rr_out, int64_t* fine_applies);
  void import_vec(const double* src, double* dst, Layout layout, int rows);
  void export_vec(const double* src, double* dst, Layout layout, int rows);
  double read_scalar(int slot);
  MeshView mesh_view() const;
  LevelView level_view(int l, bool fp32) const;

  HostMesh mesh_;
  int N_;
  RefElement fine_re_;
  hipStream_t stream_;
  bool own_stream_ = false;
  int64_t dat_stride_ = 0;  // caller's SoA stride, 0 = stride_
  int64_t stride_;
  // device mesh
  DBuf<double> Kt_, J_;
  DBuf<double> Ka_;  // 3D: the metric cell by cell, for the order-1 pass
  DBuf<int> nbr_, code_, bctype_;
  DBuf<int> nbc_;  // packed neighbour words of the element passes (empty: not packed)
  DBuf<int> nbc_cells_;  // 3D: the same words cell by cell, for the coarse CG's order-1 apply
  DBuf<double> bc_;
  bool have_bc_ = false;
  bool var_factor_ = false;
  DBuf<double> factor_;  // fine-level factor, SoA
  // over-integration of the finest order: rule, tables, coefficient at the points, traces
  bool oi_ = false;
  std::unique_ptr<OiRule> oi_rule_;
  DBuf<double> oi_C64_, oi_kv64_, oi_ks64_, oi_tr64_;
  DBuf<float> oi_C32_, oi_kv32_, oi_ks32_, oi_tr32_;
  DBuf<int> oi_qmap_;
  template <typename T>
  OiView<T> oi_view();
  void ensure_oi_tables();
  void set_oi_part(const double* k, Layout layout, bool surface);
  bool var_mm_ = false;  // per-node mass-term factor (fine level in mm_, per level in Work::mm)
  DBuf<double> mm_;
  double mm_const_ = 0;  // constant mass-term factor, 0 = none
  double inv_mass_factor_ = 1;
  DBuf<double> inv_mass_recp_;  // per-node divisor of the inverse-mass preconditioner, or empty
  DBuf<double> minv_;           // inverse reference mass matrix of the finest order
  DBuf<double> grad_xi_;        // d xi_r / d x_x per cell, SoA rows r*dim + x (built on first use)
  DBuf<double> lift_;           // M^-1 EMAT of the finest order, Np x (Nf*Nfp)
  bool has_mass_term() const { return var_mm_ || mm_const_ != 0; }

  std::vector<std::unique_ptr<Level>> levels_;
  std::unique_ptr<Amg> amg_;
  DBuf<double> blk_;  // inverse diagonal blocks (block-Jacobi)
  DBuf<double> Mfine_, Efine_;  // reference mass and EMAT of the solver order (pressure rhs)
  struct Snapshot {
    double t;
    DBuf<double> x;
  };
  std::vector<Snapshot> hist_;
  int hist_max_ = 3;
  SolverType type_ = SolverType::PMULTIGRID;
  Precision prec_ = Precision::FP64;
  ElementPass epass_ = ElementPass::AUTO;
  PMGConfig pmg_;
  bool nullspace_ = false;
  double rtol_ = 1e-8, atol_ = 1e-9;
  int max_iter_ = 500;
  bool initialised_ = false;
  bool levels_dirty_ = true;

  // outer FCG vectors (FP64, fine level)
  DBuf<double> b_, x_, r_, z_, zold_, p_, Ap_;
  // FP32 Krylov vectors of the mixed solve (its residual lives in the fine level's w32.b)
  DBuf<float> z32_, zold32_, p32_, Ap32_;
  DBuf<double> scal_;
  DBuf<double> partial_;
  double* host_scal_ = nullptr;  // pinned
  int64_t smoother_applies_ = 0;

  bool profiling_ = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_trace_, ev_apply_, ev_coarse_;
  int64_t coarse_sample_ = 0;
  std::vector<hipEvent_t> ev_pool_;
  KernelTiming timing_;
  hipEvent_t ev_take();
};

}  // namespace insp
