/*
 * ins_poisson.h - C ABI of the MI355X-native DG Poisson solve.
 *
 * Drop-in for the Poisson linear-solver interface that INS-Solver calls on
 * the (un-vendored) OP2-DG-Toolkit. The reference call sequence is
 *
 *   PoissonMatrixFreeDiag2D(mesh) / PETScPMultigrid(mesh)   src/solvers/2d/ins_solver.cpp:91-96
 *   matrix->set_bc_types(bc_types)                          src/solvers/2d/ins_solver.cpp:205-206
 *   solver->set_nullspace(b); set_tol_and_iter(r,a,it)      src/solvers/2d/ins_solver_base.cpp:154-170
 *   solver->init()                                          src/solvers/2d/ins_solver.cpp:207
 *   matrix->set_factor(1/rho)                               src/solvers/2d/mp_ins_solver.cpp:576-578
 *   solver->set_bcs(bc_data); solver->solve(rhs, ans)       src/solvers/2d/ins_solver.cpp:324-325
 *
 * and each entry point below names the call it replaces. Conventions:
 *  - return values: 0 = success / converged, 1 = not converged (the
 *    reference's `solve` returning false), < 0 = error (see INS_ERR_*);
 *    ins_poisson_last_error() describes the last error of the calling thread.
 *  - vectors are cells x Np nodal values, in one of three layouts:
 *      INS_LAYOUT_HOST_AOS    host pointer, dat[cell*Np + i] (OP2 CPU dat)
 *      INS_LAYOUT_DEVICE_SOA  device pointer, dat[i*stride + cell] with the
 *                             stride from ins_poisson_info (OP2 GPU dat with
 *                             -DDG_OP2_SOA, Makefile:40-41)
 *      INS_LAYOUT_DEVICE_AOS  device pointer, dat[cell*Np + i]
 *  - the rhs is mass-weighted (M f) as assembled by the reference callers
 *    (src/solvers/2d/ins_solver.cpp:310-319); boundary terms from
 *    ins_poisson_set_bcs are added inside the solve (viscous_solver.cpp:62-63).
 *  - `ans` is the initial guess on entry and the solution on exit.
 *  - one handle per host thread; a handle is not thread-safe.
 */
#ifndef INS_POISSON_H
#define INS_POISSON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ins_poisson* ins_poisson_t;
typedef struct ins_mesh* ins_mesh_t;

enum {
  INS_OK = 0,
  INS_NOT_CONVERGED = 1,
  INS_ERR_INVALID = -1,
  INS_ERR_HIP = -2,
  INS_ERR_BREAKDOWN = -3, /* p.Ap <= 0 or a non-finite residual */
  INS_ERR_NOMEM = -4,
  INS_ERR_INTERNAL = -5
};

enum { INS_LAYOUT_HOST_AOS = 0, INS_LAYOUT_DEVICE_SOA = 1, INS_LAYOUT_DEVICE_AOS = 2 };

/* Numbering follows LinearSolver::Solvers (src/solvers/2d/ins_solver_base.cpp:1294-1309). */
enum { INS_SOLVER_JACOBI = 1, INS_SOLVER_BLOCK_JACOBI = 2, INS_SOLVER_INV_MASS = 3,
       INS_SOLVER_PMULTIGRID = 4 };

enum { INS_PRECISION_FP64 = 0, INS_PRECISION_MIXED = 1 };

/* Boundary types per bface (include/problem_specific_2d.h:43-44). */
enum { INS_BC_DIRICHLET = 0, INS_BC_NEUMANN = 1 };

typedef struct {
  int dim;               /* 2 (triangles) or 3 (tetrahedra) */
  int64_t n_nodes;
  const double* coords;  /* n_nodes x dim (OP2 dat node_coords) */
  int64_t n_cells;
  const int* cell2nodes; /* n_cells x (dim+1) (OP2 map cell2nodes) */
} ins_mesh_desc;

typedef struct {
  int iterations;           /* outer Krylov iterations */
  int converged;
  double rel_residual;      /* true |b - A x|_2 / |b|_2 in FP64 */
  double abs_residual;
  double bnorm;
  double seconds;           /* wall time of the solve call */
  int64_t fine_applies;     /* FP64 applications of the finest operator */
  int64_t smoother_applies; /* operator applications inside the preconditioner */
} ins_solve_stats;

/* ---- meshes (stand-in for the toolkit's DGMesh2D/3D construction, src/ins2d.cpp:192) ---- */

/* Unit square, nx x ny cells split into 2 triangles, interior vertices jittered by
 * +-jitter*h (uniform, seeded). */
int ins_mesh_generate_square(int nx, int ny, double jitter, uint64_t seed, ins_mesh_t* out);
/* Unit cube, n^3 hexes split into 6 Kuhn tetrahedra, interior vertices jittered. */
int ins_mesh_generate_cube(int n, double jitter, uint64_t seed, ins_mesh_t* out);
/* Mesh from user arrays. Negatively oriented cells are re-oriented (their last two vertices
 * swapped; ins_mesh_get returns the cell2nodes in use), and the library numbers the faces and
 * boundary faces itself (ins_mesh_get_faces). */
int ins_mesh_create(const ins_mesh_desc* desc, ins_mesh_t* out);
/* The OP2 face sets of a caller's mesh (the HDF5 mesh's maps face2cells / bface2cells and dats
 * faceNum / bfaceNum, src/ins2d.cpp:192; local faces numbered 0:(v0,v1) 1:(v1,v2) 2:(v2,v0) in
 * 2D and 0:(v0,v1,v2) 1:(v0,v1,v3) 2:(v1,v2,v3) 3:(v0,v2,v3) in 3D). */
typedef struct {
  int64_t n_faces;
  const int* face2cells; /* n_faces x 2 */
  const int* faceNum;    /* n_faces x 2: local face on each side */
  int64_t n_bfaces;
  const int* bface2cells; /* n_bfaces */
  const int* bfaceNum;    /* n_bfaces */
} ins_mesh_faces_desc;
/* Mesh from user arrays that keeps the caller's face and boundary-face order, as the
 * reference's DGMesh2D/3D(filename) reads it (src/ins2d.cpp:192): bc_types and bc_data of the
 * solvers built on it are indexed by the caller's bface numbers. Every face must be a face of
 * the mesh with the same local face numbers, each given once; negatively oriented cells are
 * rejected (INS_ERR_INVALID naming the cell), as the caller's face numbers and nodal data refer
 * to its own vertex order. A face may join two boundary faces whose vertices match by a
 * translation (a periodic face). ins_mesh_renumber re-derives the face sets in the library's
 * order, and is refused on meshes with periodic faces. */
int ins_mesh_create_op2(const ins_mesh_desc* desc, const ins_mesh_faces_desc* faces,
                        ins_mesh_t* out);
/* Replaces DGMesh2D/3D(filename) (src/ins2d.cpp:192): reads an OP2 HDF5 mesh file (sets nodes,
 * cells, faces, bfaces; maps cell2nodes, face2cells, bface2cells; dats node_coords, faceNum or
 * edgeNum, bfaceNum; the files of grids/h5) and builds the mesh as ins_mesh_create_op2 does,
 * keeping the file's face and bface order. Faces that join two boundary faces whose vertices
 * match by a translation are periodic faces (the meshes with a periodicFace dat). Reads the
 * HDF5 subset OP2 writes (superblock 0/1, contiguous or compact datasets, no compression)
 * without an HDF5 library; other files give INS_ERR_INVALID. */
int ins_mesh_load_op2_hdf5(const char* path, ins_mesh_t* out);
int ins_mesh_info(ins_mesh_t m, int* dim, int64_t* n_nodes, int64_t* n_cells, int64_t* n_faces,
                  int64_t* n_bfaces);
/* Any output pointer may be NULL. */
int ins_mesh_get(ins_mesh_t m, double* coords, int* cell2nodes);
/* OP2 sets faces/bfaces: face2cells (n_faces x 2), faceNum (n_faces x 2),
 * bface2cells (n_bfaces), bfaceNum (n_bfaces). */
int ins_mesh_get_faces(ins_mesh_t m, int* face2cells, int* faceNum, int* bface2cells,
                       int* bfaceNum);
/* Nodal coordinates of the order-`order` element nodes, dim x n_cells x Np (the toolkit
 * dats x, y, z that callers evaluate right-hand sides on); host only. */
int ins_mesh_node_coords(ins_mesh_t m, int order, double* xyz);
void ins_mesh_destroy(ins_mesh_t m);
/* Renumbers the mesh's cells in place for the solver's face gathers, as OP2's op_renumber
 * reorders sets before a run (callers then lay their dats out in the new order):
 *   INS_RENUMBER_SHAPE_CLASS  group cells of the same shape class (edge directions from the
 *                             first vertex, rounded), keeping their relative order; on a
 *                             structured simplex subdivision a wave's neighbour values are
 *                             then contiguous.
 *   INS_RENUMBER_SHAPE_CLASS_WINDOWED  the same grouping within runs of 64 K cells (K shape
 *                             classes, at most 32; more classes: one run), so a cell's
 *                             neighbours stay close to it in memory.
 *   INS_RENUMBER_SHAPE_CLASS_TILED  the same grouping within space tiles (cubes or squares
 *                             holding about 64 cells of each class, in lexicographic order), so
 *                             neighbours in every direction mostly share a cell's tile.
 * perm (n_cells, may be NULL) receives new cell i = old cell perm[i]. No counterpart in the
 * reference's Poisson path. */
enum { INS_RENUMBER_SHAPE_CLASS = 1, INS_RENUMBER_SHAPE_CLASS_WINDOWED = 2, INS_RENUMBER_SHAPE_CLASS_TILED = 3 };
int ins_mesh_renumber(ins_mesh_t m, int method, int64_t* perm);

/* ---- solver handle: PoissonMatrixFreeDiag{2D,3D} + LinearSolver ---- */

/* Replaces `new PoissonMatrixFreeDiag2D(mesh)` + `new PETScPMultigrid(mesh)`
 * (src/solvers/2d/ins_solver.cpp:91-96). Copies the mesh to the device. */
int ins_poisson_create(ins_mesh_t mesh, int order, ins_poisson_t* out);
void ins_poisson_destroy(ins_poisson_t h);
int ins_poisson_info(ins_poisson_t h, int* np, int* nfp, int64_t* n_cells, int64_t* stride,
                     int64_t* n_bfaces);
/* Nodal coordinates, dim x n_cells x Np (the toolkit dats x, y, z). */
int ins_poisson_node_coords(ins_poisson_t h, double* xyz);
/* Stride of every INS_LAYOUT_DEVICE_SOA vector the handle is given afterwards, dat[i*stride +
 * cell]: the padded set size of the caller's OP2 GPU dats (-DDG_OP2_SOA with OP_AUTO_SOA,
 * Makefile:40-41, code_gen.sh:31-33; getSetSizeFromOpArg, ls_reinit_gpu.cu:13,71-76). 0 (the
 * default) = the stride reported by ins_poisson_info; otherwise at least n_cells, and then the
 * entries of padding cells (cell >= n_cells) are neither read nor written. Face dats
 * (cells x Nf*Nfp) share the stride. */
int ins_poisson_set_dat_stride(ins_poisson_t h, int64_t stride);

/* PoissonMatrix::set_bc_types(op_dat bfaces x 1) - src/solvers/2d/ins_solver.cpp:205. */
int ins_poisson_set_bc_types(ins_poisson_t h, const int* bc_types, int64_t n_bfaces);
/* FactorPoissonMatrixFreeDiag::set_factor(op_dat cells x Np) - src/solvers/2d/mp_ins_solver.cpp:577.
 * NULL restores the constant coefficient. */
int ins_poisson_set_factor(ins_poisson_t h, const double* factor, int layout);
/* Mass term of the viscous (Helmholtz) operator, A u + J M (mm o u):
 * FactorMMPoissonMatrixFree{Diag,BlockDiag}2D::set_mm_factor(op_dat cells x Np)
 * (src/solvers/2d/ins_solver.cpp:506-513, kernel fvmf_2d_mult_mm_geof.h). NULL clears. */
int ins_poisson_set_mm_factor(ins_poisson_t h, const double* mm, int layout);
/* Constant mass factor, A u + mm J M u: MMPoissonMatrixFree2D::set_factor(DG_FP)
 * (src/solvers/2d/ins_solver.cpp:518-520, kernel vmf_2d_mult_mm_geof.h). 0 = none. */
int ins_poisson_set_mm_const(ins_poisson_t h, double mm);
/* Over-integration: FactorPoissonMatrixFreeDiagOI{2D,3D}::mat_free_set_factor_oi and
 * mat_free_set_factor_surf_oi (src/solvers/2d/mp_ins_solver.cpp:597-598, selected by
 * [pressure-solve] over_int, :98-100). The coefficient at the cubature points of the cells
 * (cells x n_vol, in `layout`) and of their faces (cells x (Nfaces*n_surf), face-major, each
 * face's points in the order of ins_oi_rule); while both are set, the finest order integrates
 * the operator with them. The nodal factor of ins_poisson_set_factor still sets the penalty,
 * the boundary terms and the coarser p-multigrid levels, as the reference's caller also calls
 * set_factor (:599). NULL clears one part. */
int ins_poisson_set_factor_oi(ins_poisson_t h, const double* factor_vol, int layout);
int ins_poisson_set_factor_surf_oi(ins_poisson_t h, const double* factor_surf, int layout);
/* The cubature rules of the over-integration at order N (the library's own rules: the caller
 * evaluates its coefficient at these points in place of DGConstants::CUB2D_INTERP /
 * CUBSURF2D_INTERP, src/solvers/3d/ls_solver.cpp:202-232): n_vol volume points (reference
 * coordinates, n_vol x dim, weights summing to the reference volume) and n_surf points per face
 * (barycentrics of the face's vertices, n_surf x dim, weights summing to 2). Arrays may be NULL
 * to query the counts. */
int ins_oi_rule(int dim, int order, int* n_vol, int* n_surf, double* vol_rst, double* vol_w,
                double* surf_bary, double* surf_w);
/* Interpolation from the nodes of order N to those points: vol n_vol x Np and surf
 * (Nfaces*n_surf) x Np, row-major (either may be NULL). */
int ins_oi_interp(int dim, int order, double* vol, double* surf);
/* Scaling of the inverse-mass preconditioner (INS_SOLVER_INV_MASS), z = f (J M)^-1 r:
 * PETScInvMassSolver::setFactor / ViscousSolver2D::set_inv_mass_factor
 * (src/solvers/2d/ins_solver.cpp:521, slip_matrix/viscous_solver.cpp:49-51,347-357). */
int ins_poisson_set_inv_mass_factor(ins_poisson_t h, double f);
/* Per-node divisor instead, z_i = ((J M)^-1 r)_i / f_i: ViscousSolver2D::set_inv_mass_recp_factor
 * (slip_matrix/viscous_solver.cpp:53-55,359-366). NULL returns to the scalar factor. */
int ins_poisson_set_inv_mass_recp_factor(ins_poisson_t h, const double* f, int layout);
/* LinearSolver::set_bcs(op_dat bfaces x Nfp) - src/solvers/2d/ins_solver.cpp:324. Values are
 * given at the face nodes of each bface in the owner's face-mask order. Layout HOST_AOS or
 * DEVICE_AOS (row per bface). NULL clears. */
int ins_poisson_set_bcs(ins_poisson_t h, const double* bc_data, int layout);
/* LinearSolver::set_nullspace(bool) - src/solvers/2d/ins_solver_base.cpp:156-157. */
int ins_poisson_set_nullspace(ins_poisson_t h, int nullspace);
/* LinearSolver::set_tol_and_iter(rtol, atol, max_iter) - src/solvers/2d/ins_solver_base.cpp:170.
 * Converged when |b - A x|_2 <= max(rtol |b|_2, atol). */
int ins_poisson_set_tol_and_iter(ins_poisson_t h, double rtol, double atol, int max_iter);
/* Solver type (INS_SOLVER_*): the reference's set_solver_type enum, ins_solver_base.cpp:1294-1309. */
int ins_poisson_set_solver(ins_poisson_t h, int type);
/* INS_PRECISION_FP64, or INS_PRECISION_MIXED: FP32 V-cycle inside the FP64 flexible CG. */
int ins_poisson_set_precision(ins_poisson_t h, int precision);
/* Kernel of the element pass of the FP32 operator (the V-cycle of INS_PRECISION_MIXED and
 * ins_poisson_mult_fp32). No counterpart in the reference, whose operator has one code path:
 * a tuning choice of this library, taking effect at the next init.
 *   INS_ELEMENT_PASS_AUTO    matrix cores where they are the faster kernel (3D, order 4),
 *                            vector FMA elsewhere (default);
 *   INS_ELEMENT_PASS_VECTOR  vector FMA on every level;
 *   INS_ELEMENT_PASS_MATRIX  matrix cores on every level that has them (3D, orders 2-4). */
enum { INS_ELEMENT_PASS_AUTO = 0, INS_ELEMENT_PASS_VECTOR = 1, INS_ELEMENT_PASS_MATRIX = 2 };
int ins_poisson_set_element_pass(ins_poisson_t h, int kind);
/* p-multigrid options (the [p-multigrid] INI section: orders, pre_it, post_it, cheb_orders,
 * eigen_val_saftey_factor, cheb_min_lamda_factor, max_iter_eig_approx); per-level arrays
 * have n_levels entries and may be NULL for defaults. coarse_iters: Jacobi-CG iterations on
 * the coarsest order. */
int ins_poisson_set_pmg(ins_poisson_t h, const int* orders, int n_levels, const int* pre_it,
                        const int* post_it, const int* cheb_orders, double eig_safety,
                        double cheb_min_factor, int max_iter_eig, int coarse_iters);
/* Coarse solve of the p-multigrid V-cycle, in place of the reference's PETSc solve on
 * the assembled order-1 matrix (PoissonCoarseMatrix + set_coarse_matrix,
 * src/solvers/2d/ins_solver.cpp:92-96,205):
 *   INS_COARSE_JACOBI_CG  Jacobi-CG with coarse_iters iterations on order 1 (default);
 *   INS_COARSE_AMG        Chebyshev smoothing on order 1, then amg_cycles smoothed-aggregation
 *                         AMG V-cycles on the continuous vertex space (order 1 must be the
 *                         coarsest p-multigrid order). */
enum { INS_COARSE_JACOBI_CG = 0, INS_COARSE_AMG = 1 };
int ins_poisson_set_coarse(ins_poisson_t h, int kind, int amg_cycles);
/* The assembled coarse matrix of the AMG coarse solve, A_v = P_v^T A_1 P_v on the mesh
 * vertices (CSR, n x n). With NULL arrays only *n and *nnz are returned. Needs init with
 * INS_COARSE_AMG. */
int ins_poisson_coarse_matrix(ins_poisson_t h, int64_t* n, int64_t* nnz, int* rowptr, int* cols,
                              double* vals);
/* Number of AMG levels and the size of each (sizes may be NULL; at most 16 levels). */
int ins_poisson_amg_levels(ins_poisson_t h, int* n_levels, int* sizes);
/* LinearSolver::init() - src/solvers/2d/ins_solver.cpp:207: penalties, diagonals and
 * Chebyshev eigenvalue bounds. Called implicitly by solve when needed. */
int ins_poisson_init(ins_poisson_t h);
/* LinearSolver::solve(rhs, ans) - src/solvers/2d/ins_solver.cpp:325. */
int ins_poisson_solve(ins_poisson_t h, const double* rhs, double* ans, int layout,
                      ins_solve_stats* stats);

/* PoissonMatrix::mult(in, out) on the finest order (FP64), and its FP32 kernel variant. */
int ins_poisson_mult(ins_poisson_t h, const double* in, double* out, int layout);
int ins_poisson_mult_fp32(ins_poisson_t h, const double* in, double* out, int layout);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the FP64 apply of
 * ins_poisson_mult, calling no allocator. Its face traces (and, for layouts other than the
 * solver's own device SoA, staging copies of in and out) live in a static per-GPU scratch array
 * of INS_NOALLOC_SCRATCH_BYTES; a mesh needing more returns INS_ERR_INVALID, as does a handle not
 * yet initialised (ins_poisson_init) or a host layout. Calls that overlap on one GPU share the
 * scratch and race. */
#define INS_NOALLOC_SCRATCH_BYTES (64ll << 20)
int ins_poisson_mult_noalloc(ins_poisson_t h, const double* in, double* out, int layout);
/* Diagonal used by Jacobi (FactorViscousMatrix2D::calc_diag, factor_viscous_matrix.cpp:211-256). */
int ins_poisson_diag(ins_poisson_t h, double* diag, int layout);
/* Inverses of the per-cell Np x Np diagonal blocks used by INS_SOLVER_BLOCK_JACOBI
 * (FactorViscousMatrix2D::calc_inv_block_diag, factor_viscous_matrix.cpp:258-326), host
 * cells x Np x Np row-major (layout INS_LAYOUT_HOST_AOS only). Initialises when needed. */
int ins_poisson_inv_block_diag(ins_poisson_t h, double* out, int layout);
/* rhs += boundary terms (vmf_2d_apply_bc.h / fvmf_2d_apply_bc.h), in place. */
int ins_poisson_apply_bc(ins_poisson_t h, double* rhs, int layout);
/* out = J M in per cell (DGMesh::mass). */
int ins_poisson_mass(ins_poisson_t h, const double* in, double* out, int layout);
/* The caller's pressure rhs, fused (ins_pressure_rhs_2d.h:1-12, then op2_gemv(LIFT, dPdNOld) and
 * mesh->mass, src/solvers/2d/ins_solver.cpp:310-319): dPdN_old <- b0 dPdN + b1 dPdN_old and
 * div <- J M (-div/dt + LIFT dPdN_old) in place, ready for ins_poisson_solve. dPdN, dPdN_old are
 * cells x (Nfaces*Nfp) in face-mask order; all three in `layout` (SoA shares the stride). */
int ins_poisson_pressure_rhs(ins_poisson_t h, double b0, double b1, double dt, const double* dPdN,
                             double* dPdN_old, double* div, int layout);
/* Initial-guess extrapolation: add_to_pr_history (src/solvers/2d/ins_solver_base.cpp:1141-1155)
 * stores (t, x); initial_guess_extrapolation (src/solvers/3d/ins_solver.cpp:342-343) writes the
 * Lagrange extrapolation through the stored pairs at time t into ans (unchanged when the history
 * is empty). The history keeps the latest n (1..8, default 3) solutions on the device. */
int ins_poisson_set_history_size(ins_poisson_t h, int n);
int ins_poisson_history_push(ins_poisson_t h, double t, const double* x, int layout);
int ins_poisson_initial_guess(ins_poisson_t h, double t, double* ans, int layout);
/* Power-iteration estimate of lambda_max(D^-1 A) on p-multigrid level `level`. */
int ins_poisson_lambda_max(ins_poisson_t h, int level, double* out);

/* Device time of the finest-order SIPG kernels (HIP events on the solver's stream). */
int ins_poisson_set_profiling(ins_poisson_t h, int on);
int ins_poisson_take_timing(ins_poisson_t h, double* trace_ms, int64_t* trace_launches,
                            double* apply_ms, int64_t* apply_launches);
/* Device time of the operator applies on the coarsest p-multigrid level (order 1 with the
 * default levels: the applies of the coarse Jacobi-CG), one launch in four timed while
 * profiling is on; launches counts the timed ones. */
int ins_poisson_take_coarse_timing(ins_poisson_t h, double* apply_ms, int64_t* launches);

/* Gradient of a nodal field of the solver's order with a central flux:
 * DGMesh::grad_with_central_flux (src/solvers/2d/ins_solver.cpp:337; toolkit code not vendored:
 * strong form, p* = {p} on interior faces, p* = p- on boundary faces). dpdz is unused in 2D. */
int ins_poisson_grad_central(ins_poisson_t h, const double* p, double* dpdx, double* dpdy, double* dpdz,
                             int layout);
/* Velocity update of the projection, uTT = uT - dt grad_h p (ins_2d_proj_rhs.h:1-10, after
 * grad_with_central_flux at src/solvers/2d/ins_solver.cpp:337-339); uTz/uTTz unused in 2D. */
int ins_poisson_project_velocity(ins_poisson_t h, double dt, const double* p, const double* uTx,
                                 const double* uTy, const double* uTz, double* uTTx, double* uTTy,
                                 double* uTTz, int layout);

/* Achievable peaks of the current device, measured in about a second: streaming copy
 * bandwidth (GB/s, bytes read + written; the faster of a plain and a non-temporal copy of
 * 512 MiB, one 16-byte element per thread) and 16x16x4 matrix-core throughput in FP32 and
 * FP64 (TFLOP/s). No counterpart in the reference: the denominators of the measured
 * rooflines beside the published figures. Any output may be NULL. */
int ins_device_peaks(double* copy_gbs, double* mfma_f32_tflops, double* mfma_f64_tflops);

/* Reference-element tables of order `order` (the toolkit's DGConstants: Dr/Ds/Dt, MASS,
 * FMASK, face mass), row-major; any output may be NULL. rst: dim x Np, D: dim x Np x Np,
 * M: Np x Np, fmask: Nfaces x Nfp, Mface: Nfaces x Nfp x Nfp. Host only, no device use. */
int ins_ref_element(int dim, int order, double* rst, double* D, double* M, int* fmask,
                    double* Mface);
/* Nodal interpolation from order nc to order nf (Np(nf) x Np(nc)), the p-multigrid
 * prolongation; host only. */
int ins_ref_interp(int dim, int nf, int nc, double* P);

/* ---- batched solves: many independent problems advanced in lockstep ----
 *
 * The reference solves one problem per call; these entry points put many
 * (uneven sizes, orders 1..4, one dimension per batch) on the GPU at once. Each
 * problem runs the same flexible CG as ins_poisson_solve (viscous_solver.cpp:158-266,
 * in FP64) preconditioned by a two-level p-multigrid V-cycle (order N -> 1, the
 * reference's "orders = N,1", examples/2d/vortex/config.ini) with Chebyshev
 * smoothing and Jacobi-CG on order 1. A problem's arithmetic does not depend on
 * which other problems share the batch. Vectors are host AoS (cells x Np). */
typedef struct ins_batch* ins_batch_t;
typedef struct ins_server* ins_server_t;

typedef struct {
  int precision;          /* INS_PRECISION_FP64, or INS_PRECISION_MIXED: FP32 V-cycle */
  int cheb_order;         /* Chebyshev degree of one smoothing application (1..8) */
  int pre_it, post_it;    /* smoothing applications before / after the coarse solve */
  double eig_safety;      /* lambda_max safety factor (eigen_val_saftey_factor) */
  double cheb_min_factor; /* lower Chebyshev bound / upper (cheb_min_lamda_factor) */
  int max_iter_eig;       /* power iterations for lambda_max (max_iter_eig_approx) */
  int coarse_iters;       /* Jacobi-CG iterations on order 1 */
  int use_graph;          /* replay one captured HIP graph per lockstep iteration */
  int coarse;             /* INS_COARSE_AMG (default): Chebyshev on order 1 + vertex-space AMG;
                             INS_COARSE_JACOBI_CG: coarse_iters Jacobi-CG iterations on order 1 */
  int tile;               /* cells per tile and block: 64, 128, 256, or 0 (default) = autotuned: the
                             first batch of the process times each size on its problems and the
                             fastest is kept for the dimension and precision */
  int chunk;              /* lockstep iterations per graph replay (1..64, default 8): problems that
                             converge inside a replay are frozen on the device, and the host reads
                             the problems' state once per replay */
} ins_batch_config;

typedef struct {
  ins_mesh_t mesh;        /* read during add/submit only; may be destroyed afterwards */
  int order;              /* 1..4 */
  const int* bc_types;    /* n_bfaces INS_BC_* values, or NULL = all Dirichlet */
  const double* factor;   /* cells x Np coefficient (set_factor), or NULL = 1 */
  int nullspace;          /* set_nullspace */
  double rtol, atol;      /* set_tol_and_iter */
  int max_iter;
  const double* rhs;      /* cells x Np */
  int rhs_nodal;          /* 0: rhs is mass-weighted (M f); 1: rhs is nodal f, b = J M f */
} ins_problem;

/* Defaults: MIXED, Chebyshev 3, 1 pre / 1 post, 1.1, 0.1, 15, 40, graphs on, AMG coarse,
 * autotuned tile, 8 iterations per replay. */
void ins_batch_config_default(ins_batch_config* cfg);
/* Device bytes one problem holds while resident (admission control); n_nodes > 0 adds the
 * estimate for the AMG coarse solve of a mesh with that many vertices. */
int ins_batch_problem_bytes(int dim, int64_t n_cells, int order, int has_factor, int precision,
                            int64_t* bytes);
int ins_batch_amg_bytes(int dim, int64_t n_nodes, int64_t n_cells, int precision, int64_t* bytes);
int ins_batch_create(int dim, const ins_batch_config* cfg, ins_batch_t* out);
void ins_batch_destroy(ins_batch_t b);
/* Uploads n problems, sets them up (penalties, diagonals, Chebyshev bounds, in
 * one batched power iteration) and starts them; ids[i] names problem i. */
int ins_batch_add(ins_batch_t b, const ins_problem* probs, int n, int* ids);
/* Runs up to max_iters lockstep iterations (in replays of `chunk` iterations), stopping early
 * when none is running; *running = problems still running afterwards. */
int ins_batch_iterate(ins_batch_t b, int max_iters, int* running);
/* Ids finished since the last call, at most cap per call; *n = count returned. */
int ins_batch_take_finished(ins_batch_t b, int* ids, int cap, int* n);
/* Result of a finished problem: x (cells x Np, may be NULL) and stats (may be NULL).
 * Returns INS_OK, INS_NOT_CONVERGED or INS_ERR_BREAKDOWN as ins_poisson_solve would. */
int ins_batch_result(ins_batch_t b, int id, double* x, ins_solve_stats* stats);
/* Frees a finished problem's device memory; its id may be reused by a later add. */
int ins_batch_release(ins_batch_t b, int id);
int ins_batch_stats(ins_batch_t b, int64_t* bytes_in_use, int64_t* iterations_run, int* running);
/* Tile size in use and, when it was autotuned, the timed cost (ms) of one smoothing step for
 * tiles of 64, 128 and 256 cells (ms may be NULL; zeros when the size was given). */
int ins_batch_tuning(ins_batch_t b, int* tile, double* ms);
/* Wall seconds spent setting problems up so far (5 values): upload with penalties and
 * diagonals, Chebyshev bounds, AMG coarse space, first residual and preconditioner, and the
 * host hierarchy construction inside the AMG part. */
int ins_batch_setup_seconds(ins_batch_t b, double* t5);

/* ---- serving: continuous batching under an HBM budget ----
 * Requests are queued by ins_server_submit from any thread. A worker thread
 * admits them in arrival order while their device memory fits the budget,
 * advances all admitted solves together (one batch per dimension) and retires
 * each as it finishes, so later arrivals join the running batch. A request
 * larger than the whole budget is rejected (INS_ERR_NOMEM from wait). */
int ins_server_create(int64_t hbm_budget_bytes, const ins_batch_config* cfg2d,
                      const ins_batch_config* cfg3d, ins_server_t* out);
void ins_server_destroy(ins_server_t s);
int ins_server_submit(ins_server_t s, const ins_problem* p, int64_t* id);
int ins_server_poll(ins_server_t s, int64_t id, int* done);
/* Blocks until request id is finished and collects it (once): x (cells x Np, may be
 * NULL), stats, arrival -> result latency and arrival -> admission queueing time. */
int ins_server_wait(ins_server_t s, int64_t id, double* x, ins_solve_stats* stats,
                    double* latency_s, double* queue_s);
int ins_server_wait_all(ins_server_t s);
int ins_server_stats(ins_server_t s, int64_t* completed, int64_t* peak_bytes);

/* ---- Vector (velocity) viscous solve with slip walls ----
 * The slip-matrix family: ViscousMatrix{2D,3D} / FactorViscousMatrix{2D,3D} behind
 * Matrix2Vec / Matrix3Vec (include/slip_matrix/2d/matrix_2_vec.h:8-14), solved by
 * ViscousSolver{2D,3D} (include/slip_matrix/2d/viscous_solver.h), as the velocity step builds
 * them (src/solvers/2d/ins_solver.cpp:100-106, src/solvers/2d/mp_ins_solver.cpp:116-125).
 * The vector has dim components; every argument of type `const double* const*` (or
 * `double* const*`) is an array of dim component pointers, each a cells x Np dat in `layout`.
 * Each component is the scalar operator of ins_poisson_* (SIPG plus the mass term); bc types
 * are given per component: 0 Dirichlet, 1 Neumann, 2 slip. A boundary face whose every
 * component has type 2 is a slip wall: the normal velocity is held at zero and the tangential
 * part is free (vmf_2d_mult_avg_jump.h:7-27, vmf_3d_mult_avg_jump.h:8-40); type 2 on only some
 * components contributes nothing to them. */
typedef struct ins_viscous_vec* ins_viscous_vec_t;
enum { INS_BC_SLIP = 2 };
/* ViscousSolver2D::Preconditioners (include/slip_matrix/2d/viscous_solver.h:13-15). */
enum { INS_VEC_PRE_NONE = 0, INS_VEC_PRE_INV_MASS = 1, INS_VEC_PRE_FACTOR_INV_MASS = 2,
       INS_VEC_PRE_RECP_FACTOR_DAT_INV_MASS = 3, INS_VEC_PRE_JACOBI = 4,
       INS_VEC_PRE_BLOCK_JACOBI = 5 };
/* new ViscousMatrix2D(mesh) / FactorViscousMatrix2D(mesh) + new ViscousSolver2D(mesh)
 * (src/solvers/2d/ins_solver.cpp:100-106). */
int ins_viscous_vec_create(ins_mesh_t mesh, int order, ins_viscous_vec_t* out);
void ins_viscous_vec_destroy(ins_viscous_vec_t h);
int ins_viscous_vec_info(ins_viscous_vec_t h, int* ncomp, int* np, int* nfp, int64_t* n_cells,
                         int64_t* stride, int64_t* n_bfaces);
/* Stride of the caller's device SoA dats (as ins_poisson_set_dat_stride). */
int ins_viscous_vec_set_dat_stride(ins_viscous_vec_t h, int64_t stride);
/* Matrix2Vec::set_bc_types(u_bc_ty, v_bc_ty) (matrix_2_vec.h:11): dim arrays of n_bfaces. */
int ins_viscous_vec_set_bc_types(ins_viscous_vec_t h, const int* const* bc_types, int64_t n_bfaces);
/* FactorViscousMatrix2D::set_factor (NULL = constant 1) and set_mm_factor
 * (include/slip_matrix/2d/factor_viscous_matrix.h:20-21); ViscousMatrix2D::set_factor, the
 * constant mass factor (include/slip_matrix/2d/viscous_matrix.h:20). */
int ins_viscous_vec_set_factor(ins_viscous_vec_t h, const double* factor, int layout);
int ins_viscous_vec_set_mm_factor(ins_viscous_vec_t h, const double* mm, int layout);
int ins_viscous_vec_set_mm_const(ins_viscous_vec_t h, double mm);
/* ViscousSolver2D::set_bcs(u_bcs, v_bcs): dim arrays of bfaces x Nfp (AoS), or NULL = none. */
int ins_viscous_vec_set_bcs(ins_viscous_vec_t h, const double* const* bc_data, int layout);
/* ViscousSolver2D::set_tol_and_iter (defaults rtol 1e-8, atol 1e-9, 5000 iterations,
 * src/solvers/2d/slip_matrix/viscous_solver.cpp:21-25), set_preconditioner, set_inv_mass_factor,
 * set_inv_mass_recp_factor. */
int ins_viscous_vec_set_tol_and_iter(ins_viscous_vec_t h, double rtol, double atol, int max_iter);
int ins_viscous_vec_set_preconditioner(ins_viscous_vec_t h, int preconditioner);
int ins_viscous_vec_set_inv_mass_factor(ins_viscous_vec_t h, double f);
int ins_viscous_vec_set_inv_mass_recp_factor(ins_viscous_vec_t h, const double* f, int layout);
/* Matrix2Vec::mult(u_in, v_in, u_out, v_out) and apply_bc(u_rhs, v_rhs, u_bc, v_bc) (in place
 * on rhs, with the data of ins_viscous_vec_set_bcs). */
int ins_viscous_vec_mult(ins_viscous_vec_t h, const double* const* in, double* const* out, int layout);
int ins_viscous_vec_apply_bc(ins_viscous_vec_t h, double* const* rhs, int layout);
/* Diagonal of the Jacobi preconditioner (FactorViscousMatrix2D::calc_diag, u_diag / v_diag). */
int ins_viscous_vec_diag(ins_viscous_vec_t h, double* const* diag, int layout);
/* Inverse diagonal blocks of the block-Jacobi preconditioner (calc_inv_block_diag), host
 * dim x cells x Np x Np row-major; needs INS_VEC_PRE_BLOCK_JACOBI. */
int ins_viscous_vec_inv_block_diag(ins_viscous_vec_t h, double* out);
/* ViscousSolver2D::solve(u_rhs, v_rhs, u_ans, v_ans) (viscous_solver.cpp:59-75): boundary terms
 * added to a copy of the rhs (the caller's rhs is not modified), ans IN/OUT. 0 converged,
 * 1 not converged, <0 errors (INS_ERR_BREAKDOWN on p.Ap <= 0). */
int ins_viscous_vec_solve(ins_viscous_vec_t h, const double* const* rhs, double* const* ans, int layout,
                          ins_solve_stats* stats);

const char* ins_poisson_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* INS_POISSON_H */
