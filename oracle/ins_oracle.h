/*
 * FP64 CPU oracle for the DG Poisson solve - TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of INS-Solver's SIPG operator and Krylov solvers used
 * as the parity checker for the HIP library (and, timed, as the CPU baseline).
 * The product never links or calls it. It follows the reference's op_par_loop
 * sequence literally (face loops over jump/avg buffers, per-cell flux, mass +
 * EMAT lift, geometric transform, D^T products), so it is structurally
 * independent of the element-centric GPU formulation. Its reference element is
 * derived independently too: monomial bases with exact simplex integrals
 * instead of orthonormal Jacobi bases, and face normals from vertex geometry
 * instead of metric terms.
 *
 * Arrays are OP2 CPU layout (AoS): dat[cell*dim + i].
 */
#ifndef INS_ORACLE_H
#define INS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

oracle_t* oracle_create(int dim, int order, int64_t n_nodes, const double* coords,
                        int64_t n_cells, const int* cell2nodes);
/* As oracle_create, with the caller's face sets (interior faces may be periodic pairs). */
oracle_t* oracle_create_faces(int dim, int order, int64_t n_nodes, const double* coords,
                              int64_t n_cells, const int* cell2nodes, int64_t n_faces,
                              const int* face2cells, const int* faceNum, int64_t n_bfaces,
                              const int* bface2cells, const int* bfaceNum);
void oracle_destroy(oracle_t* o);
int oracle_np(const oracle_t* o);
int oracle_nfp(const oracle_t* o);
int64_t oracle_n_faces(const oracle_t* o);
int64_t oracle_n_bfaces(const oracle_t* o);
void oracle_bfaces(const oracle_t* o, int* bface2cells, int* bfaceNum);
void oracle_faces(const oracle_t* o, int* face2cells, int* faceNum);
void oracle_node_coords(const oracle_t* o, double* xyz); /* dim x cells x Np */
void oracle_set_bc_types(oracle_t* o, const int* bc_types);
void oracle_set_factor(oracle_t* o, const double* k); /* cells x Np, NULL = 1 */
void oracle_set_threads(int n);
/* Mass term of the viscous (Helmholtz) operator on the finest order: per-node factor
 * (cells x Np, NULL = none) or a constant (0 = none). */
void oracle_set_mm_factor(oracle_t* o, const double* mm);
void oracle_set_mm_const(oracle_t* o, double mm);
/* Flexible PCG preconditioned by the inverse mass matrix, scaled by inv_mass_factor or
 * divided by recp_factor (cells x Np) when that is not NULL. */
int oracle_solve_inv_mass(oracle_t* o, const double* rhs, double* x, double rtol, double atol,
                          int max_iter, double inv_mass_factor, const double* recp_factor,
                          int* iters, double* relres);
/* Order used in the penalty 2(N+1)(N+2)fscale (default: the element order). */
void oracle_set_penalty_order(oracle_t* o, int n);

/* out = A in */
void oracle_mult(oracle_t* o, const double* in, double* out);
/* Jacobi diagonal as the reference computes it (fvmf_*_op1/op2/bop_diag); with
 * consistent != 0 the factor of the symmetric face term is taken at the column
 * node, which makes it the exact diagonal of oracle_mult for any factor. */
void oracle_diag(oracle_t* o, int consistent, double* out);
/* rhs += boundary terms (fvmf_2d_apply_bc.h / vmf_2d_apply_bc.h), bc is bfaces x Nfp */
void oracle_apply_bc(oracle_t* o, const double* bc, double* rhs);
void oracle_mass(oracle_t* o, const double* in, double* out);
/* Gradient with central flux of a nodal field (cells x Np each component; gz unused in 2D). */
void oracle_grad_central(oracle_t* o, const double* p, double* gx, double* gy, double* gz);
/* Mass-weighted squared L2 norm, sum_e 1^T J M (e o e) (l2_error_vortex.cpp:25-44). */
double oracle_l2sq(oracle_t* o, const double* e);

/* Vector (velocity) operator with slip walls (Matrix2Vec / Matrix3Vec: ViscousMatrix2D,
 * FactorViscousMatrix2D and their 3D twins). nc = dim components; vectors nc x cells x Np;
 * bct nc x bfaces: 0 Dirichlet, 1 Neumann, 2 slip (a wall where every component is 2). */
void oracle_mult_vec(oracle_t* o, int nc, const int* bct, const double* in, double* out);
void oracle_diag_vec(oracle_t* o, int nc, const int* bct, int consistent, double* out);
void oracle_apply_bc_vec(oracle_t* o, int nc, const int* bct, const double* bc, double* rhs);
/* ViscousSolver2D::solve: rhs plus boundary terms of bc (nc x bfaces x Nfp, or NULL), then PCG
 * with precond 0 none, 1 inverse mass, 2 inverse mass times inv_mass_factor, 3 inverse mass
 * divided by recp_factor (cells x Np), 4 Jacobi. x IN/OUT; returns as oracle_solve_jacobi. */
int oracle_solve_vec(oracle_t* o, int nc, const int* bct, const double* bc, const double* rhs, double* x,
                     double rtol, double atol, int max_iter, int precond, double inv_mass_factor,
                     const double* recp_factor, int* iters, double* relres);

/* Over-integrated variable-coefficient operator (FactorPoissonMatrixFreeDiagOI): out = A in with
 * every integral taken by the given rules (nv volume points in reference coordinates and
 * weights; ns points per face as barycentrics of the face's vertices and weights summing to 2)
 * and the coefficient at the points: kv cells x nv, ks cells x Nf x ns. The penalty takes the
 * nodal factor of oracle_set_factor; the mass term is added as in oracle_mult. */
void oracle_mult_oi(oracle_t* o, int nv, const double* vrst, const double* vw, int ns, const double* sbary,
                    const double* sw, const double* kv, const double* ks, const double* in, double* out);

/* Flexible PCG with Jacobi preconditioner (viscous_solver.cpp:158-266); converged
 * when |b - A x| <= max(rtol |b|, atol) for the true residual. x is IN/OUT.
 * Returns 0 converged, 1 not converged, -3 breakdown. */
int oracle_solve_jacobi(oracle_t* o, const double* rhs, double* x, double rtol, double atol,
                        int max_iter, int nullspace, int* iters, double* relres);

/* Flexible CG preconditioned by a p-multigrid V-cycle: Chebyshev(cheb_orders[l])
 * smoothing with Jacobi, bounds [min_factor*safety*lmax, safety*lmax] from
 * max_iter_eig power iterations, Jacobi-CG with coarse_iters iterations on the
 * coarsest order, nodal interpolation between orders and its transpose. */
int oracle_solve_pmg(oracle_t* o, const int* orders, int n_levels, const int* pre_it,
                     const int* post_it, const int* cheb_orders, double eig_safety,
                     double cheb_min_factor, int max_iter_eig, int coarse_iters,
                     const double* rhs, double* x, double rtol, double atol, int max_iter,
                     int nullspace, int* iters, double* relres, double* lmax_out);

/* Reference-element matrices for tests (row-major). Any pointer may be NULL.
 * D is dim x Np x Np, Mface is Nfaces x Nfp x Nfp, rst is dim x Np. */
int oracle_ref_element(int dim, int order, double* rst, double* D, double* M, int* fmask,
                       double* Mface);
/* Interpolation matrix from order nc to order nf on the reference simplex (Npf x Npc). */
int oracle_interp(int dim, int nf, int nc, double* P);

#ifdef __cplusplus
}
#endif

#endif
