// The vector (velocity) viscous solve with slip walls: the reference's slip-matrix family,
// ViscousMatrix{2D,3D} / FactorViscousMatrix{2D,3D} behind Matrix2Vec / Matrix3Vec
// (include/slip_matrix/2d/matrix_2_vec.h:8-14), solved by ViscousSolver{2D,3D}
// (include/slip_matrix/2d/viscous_solver.h). Each component is the scalar SIPG operator
// plus its mass term; boundary faces carry a bc type per component (0 Dirichlet, 1 Neumann,
// 2 slip), and a face whose every component is of type 2 couples the components through the
// normal projection n n^T (vmf_2d_mult_avg_jump.h:7-27). A face of type 2 on only some
// components contributes nothing to them, in the operator and in the boundary terms.
//
// The operator, penalties, factors and mass terms of a PoissonSolver of the same mesh and
// order are shared; vectors are d stacked SoA slabs (component c at c * Np * stride).
#pragma once

#include "solver.h"

namespace insp {

// ViscousSolver2D::Preconditioners (include/slip_matrix/2d/viscous_solver.h:13-15).
enum class VecPrecond {
  NONE = 0,
  INV_MASS = 1,
  FACTOR_INV_MASS = 2,
  RECP_FACTOR_DAT_INV_MASS = 3,
  JACOBI = 4,
  BLOCK_JACOBI = 5
};

class VecViscousSolver {
 public:
  VecViscousSolver(HostMesh mesh, int order);
  ~VecViscousSolver();
  VecViscousSolver(const VecViscousSolver&) = delete;
  VecViscousSolver& operator=(const VecViscousSolver&) = delete;

  int ncomp() const { return ncomp_; }
  // The scalar solver whose operator, factor, mass term, inverse-mass factors and dat stride
  // the vector operator uses (set_factor, set_mm_factor, set_mm_const, set_inv_mass_*,
  // set_dat_stride, node_coords, info).
  PoissonSolver& base() { return base_; }
  void touch() { initialised_ = false; }

  // Matrix2Vec::set_bc_types(u_bc_ty, v_bc_ty[, w_bc_ty]): one array of n_bfaces per component.
  void set_bc_types(const int* const types[], int64_t n);
  // ViscousSolver2D::set_bcs(u_bcs, v_bcs): bfaces x Nfp per component (null: no BC data).
  void set_bcs(const double* const bc[], Layout layout);
  void set_tol_and_iter(double rtol, double atol, int max_iter);
  void set_preconditioner(VecPrecond p) { prec_ = p; initialised_ = false; }
  void init();

  void mult(const double* const in[], double* const out[], Layout layout);
  void diag(double* const out[], Layout layout);
  // Inverse diagonal blocks of the block-Jacobi preconditioner, host, comps x cells x Np x Np.
  void inv_block_diag(double* out);
  void apply_bc(double* const rhs[], Layout layout);
  int solve(const double* const rhs[], double* const ans[], Layout layout, SolveStats* stats);

 private:
  void apply(const double* u, double* out, double* partial);
  void apply_one(int c, const double* u, double* out_c, double* partial);
  void precondition(const double* r, double* z);
  void build_block_jacobi();
  void import(const double* const src[], double* dst, Layout layout);
  void export_to(const double* src, double* const dst[], Layout layout);
  double read_scalars();

  PoissonSolver base_;
  int ncomp_ = 0;
  std::vector<int> types_;  // ncomp x n_bfaces as given
  DBuf<int> eff_;           // ncomp x n_bfaces: 0 Dirichlet, 1 Neumann, 2 slip, 3 nothing
  DBuf<double> bn_;         // n_bfaces x dim outward unit normals
  DBuf<double> bc_;         // ncomp x n_bfaces x Nfp
  bool have_bc_ = false;
  VecPrecond prec_ = VecPrecond::NONE;
  // ViscousSolver2D's defaults (src/solvers/2d/slip_matrix/viscous_solver.cpp:21-25)
  double rtol_ = 1e-8, atol_ = 1e-9;
  int max_iter_ = 5000;
  bool initialised_ = false;
  DBuf<double> b_, x_, r_, z_, zold_, p_, Ap_, tr_, diag_, inv_, blk_, minv_, partial_, scal_;
  double* host_scal_ = nullptr;
};

}  // namespace insp
