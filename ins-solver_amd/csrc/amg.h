// Coarse space below order 1: the continuous piecewise-linear (vertex) space and
// a smoothed-aggregation algebraic multigrid hierarchy on it.
//
// The order-1 DG space contains the continuous P1 functions; for them every
// interior-face term of the SIPG form vanishes, so the Galerkin operator
// A_v = P_v^T A_1 P_v (P_v: vertex value -> the DG nodes at that vertex) is the
// P1 stiffness form with the Nitsche boundary terms of A_1 (symmetric for a
// constant coefficient; the AMG is built from its symmetric part). A_v is assembled
// exactly on the device by probing A_1 with vertex colour classes (a distance-2
// colouring of the vertex graph keeps the columns of one class apart), so it
// takes the coefficient, penalty and boundary types from the operator itself.
// This is the assembled coarse matrix of the reference's p-multigrid
// (PoissonCoarseMatrix / set_coarse_matrix, src/solvers/2d/ins_solver.cpp:92-96,205),
// taken one step further down, and the AMG stands in for the PETSc coarse solve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace insp {

struct HostMesh;

struct HostCsr {
  int n = 0, m = 0;  // rows, columns
  std::vector<int> ptr, col;
  std::vector<double> val;
  int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

HostCsr csr_transpose(const HostCsr& A);
HostCsr csr_multiply(const HostCsr& A, const HostCsr& B);
void csr_matvec(const HostCsr& A, const double* x, double* y);

// Vertex graph of a simplex mesh (vertices sharing a cell, diagonal included, sorted).
HostCsr vertex_graph(const HostMesh& m);
// Greedy distance-2 colouring: two vertices of one colour share no neighbour.
std::vector<int> distance2_colouring(const HostCsr& g, int* n_colours);

struct AmgOptions {
  double theta = 0.02;      // strength threshold |a_ij| >= theta sqrt(a_ii a_jj)
  double omega = 4.0 / 3.0; // prolongator smoothing weight / lambda_max(D^-1 A)
  int max_coarse = 256;     // dense solve at or below this size
  int max_levels = 12;
  int cheb_order = 2;       // Chebyshev degree of one smoothing step on each level
};

struct AmgHostLevel {
  HostCsr A, P, R;           // operator, prolongation to this level from the next, R = P^T
  std::vector<double> diag;
  double lmax = 1.0;         // lambda_max(D^-1 A), power iteration
};

struct AmgHierarchy {
  std::vector<AmgHostLevel> levels;  // levels[0] = vertex space
  int n_coarse = 0;
  std::vector<double> coarse_inv;    // dense inverse of the coarsest operator (row-major)
};

// Builds the hierarchy from a symmetric positive (semi-)definite A. With
// `singular` the coarsest inverse is taken of A + (mean diag / n) 1 1^T, which
// acts as the pseudo-inverse on mean-free right-hand sides.
AmgHierarchy amg_setup(HostCsr A, const AmgOptions& opt, bool singular);

template <typename T>
struct DevCsr {
  int n = 0, m = 0;
  const int* ptr = nullptr;
  const int* col = nullptr;
  const T* val = nullptr;
  int lanes = 1;  // lanes per row of the product (lanes_per_row), 1: one row per lane
};

// Lanes that share one row of a CSR product: about two nonzeros per lane, 2..32.
inline int lanes_per_row(int64_t nnz, int n) {
  const double avg = n > 0 ? (double)nnz / n : 0.0;
  return avg < 6 ? 2 : avg < 12 ? 4 : avg < 24 ? 8 : avg < 48 ? 16 : 32;
}

namespace amg {
// y = A x (mode 0), y += A x (mode 1), y = b - A x (mode 2).
template <typename T>
void spmv(const DevCsr<T>& A, const T* x, T* y, const T* b, int mode, hipStream_t s);
// y = M x for a dense n x n M.
template <typename T>
void dense_matvec(int n, const T* M, const T* x, T* y, hipStream_t s);
// Vertex restriction: bv[v] = sum of r over the DG order-1 entries at v (gather list).
template <typename T>
void vertex_restrict(int nv, const int* vptr, const int* vlist, const T* r, T* bv, hipStream_t s);
// Vertex prolongation: x[i*stride+e] (+)= xv[cv[i*stride+e]] for e < E.
template <typename T>
void vertex_prolong(int rows, int64_t E, int64_t stride, const int* cv, const T* xv, T* x, bool add,
                    hipStream_t s);
// Probe vector: x[v] = (colour[v] == c).
void probe_fill(int nv, const int* colour, int c, double* x, hipStream_t s);
// val[k] = y[r] for the entries k of row r whose column has colour c.
void probe_extract(int nv, const int* ptr, const int* col, const int* colour, int c, const double* y,
                   double* val, hipStream_t s);
}  // namespace amg

}  // namespace insp
