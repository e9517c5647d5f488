// Layout of the per-order operator-constant buffer and of the index table
// consumed by the SIPG kernels. Shared by host (buffer construction) and
// device (compile-time offsets).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace insp {

struct OpLayout {
  int dim, N, Np, Nfp, Nf, NK, nperm;
  // Real-valued constants (stored in the kernel's working precision):
  //  D[s][j][m]     reference derivative matrices, row j contiguous
  //  Bt[r][j][i]    (Dr^T M)[i][j]: column j of Dr^T M contiguous
  //  MfT[f][j][i]   face mass Mface_f[i][j], column j contiguous
  //  Wt[f][r][j][i] (Dr^T EMAT_f)[i][j], column j contiguous
  //  M[i][j]        reference mass matrix
  int off_D, off_Bt, off_MfT, off_Wt, off_M, size;
  // Integer table:
  //  fmask[f][j]
  //  nbr[f][fn][p][j] = (neighbour volume node << 6) | (fn*Nfp + q)
  //  finv[f][i]       position of node i in fmask_f, or -1
  int tab_fmask, tab_nbr, tab_finv, tab_size;
};

__host__ __device__ constexpr OpLayout op_layout(int dim, int N) {
  OpLayout L{};
  L.dim = dim;
  L.N = N;
  L.Np = dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
  L.Nfp = dim == 2 ? N + 1 : (N + 1) * (N + 2) / 2;
  L.Nf = dim + 1;
  L.NK = dim == 2 ? 3 : 6;
  L.nperm = dim == 2 ? 2 : 6;
  L.off_D = 0;
  L.off_Bt = L.off_D + dim * L.Np * L.Np;
  L.off_MfT = L.off_Bt + dim * L.Np * L.Np;
  L.off_Wt = L.off_MfT + L.Nf * L.Nfp * L.Nfp;
  L.off_M = L.off_Wt + L.Nf * dim * L.Nfp * L.Np;
  L.size = L.off_M + L.Np * L.Np;
  L.tab_fmask = 0;
  L.tab_nbr = L.tab_fmask + L.Nf * L.Nfp;
  L.tab_finv = L.tab_nbr + L.Nf * L.Nf * L.nperm * L.Nfp;
  L.tab_size = L.tab_finv + L.Nf * L.Np;
  return L;
}

// Order 1: the nodes are the reference vertices and the basis is phi_0 = -(1 + sum_k xi_k)/2,
// phi_{k+1} = (1 + xi_k)/2, so every derivative is a constant and the order-1 kernels use these
// values as compile-time constants (the table builder checks D, M and Mface against them).
//   d phi_i / d xi_r
__host__ __device__ constexpr double p1_dphi(int r, int i) {
  return i == 0 ? -0.5 : (i == r + 1 ? 0.5 : 0.0);
}
//   column sum of the reference mass matrix (the integral of phi_j): area 2 / 3, volume 4/3 / 4
__host__ __device__ constexpr double p1_mass_colsum(int dim) { return dim == 2 ? 2.0 / 3.0 : 1.0 / 3.0; }
//   face mass in face coordinates: c (1 + delta_ij), c = length 2 / 6 or area 2 / 12
__host__ __device__ constexpr double p1_face_mass_c(int dim) { return dim == 2 ? 1.0 / 3.0 : 1.0 / 6.0; }

// Signed face-direction coefficients: the outward normal of face f is
// n_f = grad(sum_r alpha_f[r] xi_r) / |.|, xi = (r, s[, t]).
__host__ __device__ constexpr double face_alpha(int dim, int f, int r) {
  if (dim == 2) {
    const double a[3][2] = {{0, -1}, {1, 1}, {-1, 0}};
    return a[f][r];
  }
  const double a[4][3] = {{0, 0, -1}, {0, -1, 0}, {1, 1, 1}, {-1, 0, 0}};
  return a[f][r];
}

// Index of the compact symmetric metric entry (r,s): 2D {rr, rs, ss},
// 3D {rr, rs, rt, ss, st, tt}.
__host__ __device__ constexpr int kidx(int dim, int r, int s) {
  if (r > s) { const int t = r; r = s; s = t; }
  if (dim == 2) return r == 0 ? s : 2;
  return r == 0 ? s : (r == 1 ? 2 + s : 5);
}

// Node map of the order-1 faces as 2-bit fields: the neighbour's node at face node j of face f,
// for the neighbour code c = fn * nperm + permutation, is (w[f][j] >> 2c) & 3 (c < 24 in 3D; in
// 3D every row f is the same).
struct P1NodeBits {
  unsigned long long w[4][3];
};

// Compile-time face masks (combinatorial: which warp & blend nodes lie on
// which face, in volume node order). Must agree with the numerical masks of
// the reference element; the host checks this at start-up.
struct FmaskTable {
  int v[4 * 15];
};

__host__ __device__ constexpr FmaskTable make_fmask(int dim, int N) {
  FmaskTable t{};
  const int Nfp = dim == 2 ? N + 1 : (N + 1) * (N + 2) / 2;
  int cnt[4] = {0, 0, 0, 0};
  int p = 0;
  if (dim == 2) {
    for (int n = 0; n <= N; ++n)
      for (int m = 0; m <= N - n; ++m) {
        const bool on[3] = {n == 0, m + n == N, m == 0};
        for (int f = 0; f < 3; ++f)
          if (on[f]) t.v[f * Nfp + cnt[f]++] = p;
        ++p;
      }
  } else {
    for (int n = 0; n <= N; ++n)
      for (int m = 0; m <= N - n; ++m)
        for (int q = 0; q <= N - n - m; ++q) {
          const bool on[4] = {n == 0, m == 0, q + m + n == N, q == 0};
          for (int f = 0; f < 4; ++f)
            if (on[f]) t.v[f * Nfp + cnt[f]++] = p;
          ++p;
        }
  }
  return t;
}

}  // namespace insp
