// Block-Jacobi preconditioner: per-cell Np x Np diagonal blocks of the SIPG
// operator and their inverses (the reference's calc_inv_block_diag,
// src/solvers/2d/slip_matrix/factor_viscous_matrix.cpp:258-326, which inverts
// on the host; here the blocks are probed from the device operator and
// inverted on the device).
//
// Blocks are stored SoA over cells, entry (i, j) of cell e at
// blk[(i*Np + j)*stride + e], so every kernel below reads and writes coalesced
// with one thread per cell.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.h"

namespace insp {

namespace {

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("block-Jacobi kernel launch failed: ") + what + ": " +
                             hipGetErrorString(e));
}

constexpr int kB = 256;
inline unsigned cells_grid(int64_t E) { return (unsigned)((E + kB - 1) / kB); }

// u = unit vector at node j of every cell of colour c (other cells and padding 0).
__global__ void __launch_bounds__(kB)
k_probe_unit(int rows, int64_t E, int64_t stride, const int* __restrict__ colour, int c, int j,
             double* __restrict__ u) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= stride) return;
  const bool on = e < E && colour[e] == c;
  for (int i = 0; i < rows; ++i) u[i * stride + e] = (on && i == j) ? 1.0 : 0.0;
}

// Column j of the block of every cell of colour c: blk(:, j) = y(:, e).
__global__ void __launch_bounds__(kB)
k_extract_column(int rows, int64_t E, int64_t stride, const int* __restrict__ colour, int c, int j,
                 const double* __restrict__ y, double* __restrict__ blk) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= E || colour[e] != c) return;
  for (int i = 0; i < rows; ++i) blk[((int64_t)i * rows + j) * stride + e] = y[i * stride + e];
}

// In-place Gauss-Jordan inversion of each cell's SPD block (no pivoting is
// needed for symmetric positive definite matrices).
__global__ void __launch_bounds__(kB)
k_invert_blocks(int rows, int64_t E, int64_t stride, double* __restrict__ blk) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= E) return;
  auto A = [&](int i, int j) -> double& { return blk[((int64_t)i * rows + j) * stride + e]; };
  for (int k = 0; k < rows; ++k) {
    const double piv = 1.0 / A(k, k);
    A(k, k) = 1.0;
    for (int j = 0; j < rows; ++j) A(k, j) *= piv;
    for (int i = 0; i < rows; ++i) {
      if (i == k) continue;
      const double f = A(i, k);
      A(i, k) = 0.0;
      for (int j = 0; j < rows; ++j) A(i, j) = fma(-f, A(k, j), A(i, j));
    }
  }
}

// z = B^-1 r per cell.
template <int NP>
__global__ void __launch_bounds__(kB)
k_block_apply(int64_t E, int64_t stride, const double* __restrict__ blk, const double* __restrict__ r,
              double* __restrict__ z) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= E) return;
  double rv[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) rv[j] = r[j * stride + e];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NP; ++j) s = fma(blk[((int64_t)i * NP + j) * stride + e], rv[j], s);
    z[i * stride + e] = s;
  }
}

}  // namespace

void block_probe_unit(int rows, int64_t E, int64_t stride, const int* colour, int c, int j, double* u,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_probe_unit, dim3(cells_grid(stride)), dim3(kB), 0, s, rows, E, stride, colour, c,
                     j, u);
  check("k_probe_unit");
}

void block_extract_column(int rows, int64_t E, int64_t stride, const int* colour, int c, int j,
                          const double* y, double* blk, hipStream_t s) {
  hipLaunchKernelGGL(k_extract_column, dim3(cells_grid(E)), dim3(kB), 0, s, rows, E, stride, colour, c,
                     j, y, blk);
  check("k_extract_column");
}

void block_invert(int rows, int64_t E, int64_t stride, double* blk, hipStream_t s) {
  hipLaunchKernelGGL(k_invert_blocks, dim3(cells_grid(E)), dim3(kB), 0, s, rows, E, stride, blk);
  check("k_invert_blocks");
}

void block_apply(int rows, int64_t E, int64_t stride, const double* blk, const double* r, double* z,
                 hipStream_t s) {
  const dim3 g(cells_grid(E)), b(kB);
  switch (rows) {
    case 3: hipLaunchKernelGGL((k_block_apply<3>), g, b, 0, s, E, stride, blk, r, z); break;
    case 4: hipLaunchKernelGGL((k_block_apply<4>), g, b, 0, s, E, stride, blk, r, z); break;
    case 6: hipLaunchKernelGGL((k_block_apply<6>), g, b, 0, s, E, stride, blk, r, z); break;
    case 10: hipLaunchKernelGGL((k_block_apply<10>), g, b, 0, s, E, stride, blk, r, z); break;
    case 15: hipLaunchKernelGGL((k_block_apply<15>), g, b, 0, s, E, stride, blk, r, z); break;
    case 20: hipLaunchKernelGGL((k_block_apply<20>), g, b, 0, s, E, stride, blk, r, z); break;
    case 35: hipLaunchKernelGGL((k_block_apply<35>), g, b, 0, s, E, stride, blk, r, z); break;
    default: throw std::runtime_error("block_apply: unsupported block size");
  }
  check("k_block_apply");
}

}  // namespace insp
