// Over-integration of the variable-coefficient operator: the reference's
// FactorPoissonMatrixFreeDiagOI{2D,3D} (src/solvers/2d/mp_ins_solver.cpp:98-100,580-600), whose
// coefficient is given at cubature points of the cells (mat_free_set_factor_oi, cells x
// DG_CUB_NP) and of their faces (mat_free_set_factor_surf_oi, cells x Nf*DG_CUB_SURF_NP)
// instead of at the nodes.
//
// The operator is the symmetric interior penalty form integrated with those rules:
//   a(u, v) = sum_K sum_q w_q J k_q grad u . grad v (x_q)
//           + sum_F sum_q w_q sJ [ tau/2 [u] v - {k grad u}.n v - k^- [u]/2 n.grad v^- ]
// at the face points from each side (Dirichlet faces: [u] = 2u, {k grad u} = k grad u; Neumann
// faces: none), with the penalty tau of the nodal factor (fvmf_2d_calc_tau_faces.h:9).
//
// Rules (all exact for polynomials of degree 2N at least):
//  * triangle volume (2D) and triangle faces (3D): fully symmetric rules of degree 2N with
//    3, 6, 12, 16 points at N = 1..4 (orbit parameters of D. A. Dunavant, "High degree efficient
//    symmetrical Gaussian quadrature rules for the triangle", IJNME 21, 1985, refined here to
//    double precision by Newton on the moment equations); the reference's counts.
//  * edges (2D faces): Gauss-Legendre with 2N+1 points; the reference's counts.
//  * tetrahedron volume (3D): the conical (collapsed) product of Gauss-Legendre rules with
//    N+2 points per direction (degree 2N+1).
// Face points are given in barycentric coordinates of the face's vertices; the rules are
// symmetric, so the neighbour finds the same point by permuting them.
#pragma once

#include <vector>

#include "refelem.h"

namespace insp {

struct OiRule {
  int dim = 0, N = 0;
  int nv = 0, ns = 0;          // volume points, points per face
  std::vector<double> vx;      // nv x dim reference coordinates
  std::vector<double> vw;      // nv weights (sum: the reference volume, 2 or 4/3)
  std::vector<double> sb;      // ns x dim barycentrics of the face vertices
  std::vector<double> sw;      // ns weights (sum: 2, the reference face measure)
  std::vector<int> qmap;       // nperm x ns: the neighbour's point under face permutation p
};

OiRule make_oi_rule(int dim, int N);

// Operator tables of one order in double, laid out for the kernels (OiLayout).
struct OiLayout {
  int Np, Nf, nv, ns, dim;
  int off_GV, off_wv, off_IS, off_GS, off_ws, size;
};
OiLayout oi_layout(int dim, int N, const OiRule& q);
// GV[s][q][m] = d phi_m / d xi_s at volume point q; wv[q]; IS[f][q][m] = phi_m at face point q
// of face f; GS[f][s][q][m] = d phi_m / d xi_s there; ws[q].
std::vector<double> make_oi_tables(const RefElement& re, const OiRule& q);
// Interpolation from the nodes to the points (nv x Np, Nf*ns x Np) and the points' reference
// coordinates (volume nv x dim, faces Nf*ns x dim).
void oi_interp(const RefElement& re, const OiRule& q, std::vector<double>& Iv, std::vector<double>& Is,
               std::vector<double>& xv, std::vector<double>& xs);

}  // namespace insp
