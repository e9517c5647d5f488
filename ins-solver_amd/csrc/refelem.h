// Reference simplex elements for the nodal DG Poisson operator (triangles and
// tetrahedra, orders 1..kMaxOrder).
//
// Conventions (shared by every kernel and by the C ABI):
//  * Reference triangle vertices  v0=(-1,-1) v1=(1,-1) v2=(-1,1);
//    faces f0=(v0,v1) [s=-1], f1=(v1,v2) [r+s=0], f2=(v2,v0) [r=-1].
//  * Reference tetrahedron vertices v0=(-1,-1,-1) v1=(1,-1,-1) v2=(-1,1,-1)
//    v3=(-1,-1,1); faces f0=(v0,v1,v2) [t=-1], f1=(v0,v1,v3) [s=-1],
//    f2=(v1,v2,v3) [r+s+t=-1], f3=(v0,v2,v3) [r=-1].
//  * Nodes are Warburton's warp & blend nodes (Hesthaven & Warburton,
//    "Nodal Discontinuous Galerkin Methods", 2008, sections 6.1 and 10.1),
//    enumerated with r fastest, then s, then t.
//  * The face mask of face f lists the volume nodes lying on f, in volume
//    node order. Face-local buffers are indexed f*Nfp + j in that order
//    (the OP2 FMASK convention, e.g. src/kernels/vmf_2d_mult_avg_jump.h:33).
//  * All matrices are dense row-major: A[i*ncols + j].
#pragma once

#include <vector>

namespace insp {

constexpr int kMaxOrder = 4;

struct RefElement {
  int dim = 0, N = 0, Np = 0, Nfp = 0, Nfaces = 0;
  std::vector<double> r, s, t;      // node coordinates (t empty in 2D)
  std::vector<double> V, invV;      // orthonormal-basis Vandermonde
  std::vector<double> D[3];         // Dr, Ds, Dt (Np x Np)
  std::vector<double> M, invM;      // mass matrix on the reference element
  std::vector<int> fmask;           // Nfaces x Nfp
  std::vector<double> Mface;        // Nfaces x Nfp x Nfp face mass (face coords)
  std::vector<double> EMAT;         // Np x (Nfaces*Nfp): sum of face masses
  std::vector<double> face_bary;    // Nfaces x Nfp x dim barycentrics wrt face vertices
  // Node matching across a shared face.
  //   nbr_face_node[((f*Nfaces + fn)*nperm + p)*Nfp + j] = q
  // means: face node j of my face f coincides with face node q of the
  // neighbour's face fn when the neighbour's face vertex k is my face vertex
  // perms[p][k].
  int nperm = 0;
  std::vector<int> perms;           // nperm x dim
  std::vector<int> nbr_face_node;
};

// Vertex indices of each reference face, dim entries per face.
const int* face_vertices(int dim);

RefElement make_ref_element(int dim, int N);
// The order-1 neighbour node map of the operator table in bit fields (dg_layout.h P1NodeBits),
// built once per dimension.
const struct P1NodeBits& p1_node_bits(int dim);

// Nodal interpolation from order `from` to order `to` on the same reference
// simplex: returns to.Np x from.Np.
std::vector<double> interp_matrix(const RefElement& to, const RefElement& from);

// Values (n x Np) and, when G is given, reference gradients (G[d], n x Np) of the nodal basis
// at n points (reference coordinates, n x dim).
void basis_at(const RefElement& re, const std::vector<double>& pts, std::vector<double>& I,
              std::vector<double>* G);

// Small dense helpers (row-major).
std::vector<double> mat_inverse(const std::vector<double>& A, int n);
std::vector<double> mat_mul(const std::vector<double>& A, const std::vector<double>& B,
                            int m, int k, int n);
std::vector<double> mat_transpose(const std::vector<double>& A, int m, int n);

}  // namespace insp
