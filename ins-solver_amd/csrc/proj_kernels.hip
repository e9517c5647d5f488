// Pressure projection around the solve: the gradient of the pressure with a central
// flux and the velocity update of the projection step.
//
//   grad_h p = sum_r grad(xi_r) D_r p + LIFT_f (fscale_f n_f (p* - p-)),  p* = (p- + p+)/2
//   uTT = uT - dt grad_h p                                  (ins_2d_proj_rhs.h:1-10)
//
// The gradient is the toolkit's DGMesh::grad_with_central_flux (called at
// src/solvers/2d/ins_solver.cpp:337, code not vendored): the textbook strong form with a
// central flux; on boundary faces p* = p- (no face term). fscale_f n_f = sum_r alpha_fr
// grad(xi_r) for a straight-sided simplex, so only grad(xi) is stored per cell. One
// thread per cell; the neighbour's face values are gathered by index and permutation.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "dg_layout.h"
#include "kernels.h"
#include "sipg_cell.h"

namespace insp {

namespace {

template <int DIM, int N>
__global__ void __launch_bounds__(256)
k_grad_central(int64_t E, int64_t stride, const double* __restrict__ G, const int* __restrict__ nbr,
               const int* __restrict__ code, const double* __restrict__ C, const int* __restrict__ tab,
               const double* __restrict__ lift, const double* __restrict__ p, double dt,
               const double* __restrict__ u0, const double* __restrict__ u1,
               const double* __restrict__ u2, double* __restrict__ o0, double* __restrict__ o1,
               double* __restrict__ o2) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf, NF = Nf * Nfp;
  constexpr FmaskTable FM = make_fmask(DIM, N);
  const cptr<double> Cc = as_const(C);
  const cptr<double> Lc = as_const(lift);
  const cptr<int> tb = as_const(tab);
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double pp[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) pp[m] = p[m * stride + e];
  double g[DIM][DIM];  // g[r][x] = d xi_r / d x_x
#pragma unroll
  for (int r = 0; r < DIM; ++r)
#pragma unroll
    for (int x = 0; x < DIM; ++x) g[r][x] = G[(r * DIM + x) * stride + e];
  double* o[3] = {o0, o1, o2};
  const double* uu[3] = {u0, u1, u2};
  // one component at a time (a rolled loop) keeps the live registers near 2 Np
#pragma unroll 1
  for (int x = 0; x < DIM; ++x) {
    double gx[DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r) gx[r] = x == 0 ? g[r][0] : (x == 1 ? g[r][1 % DIM] : g[r][DIM - 1]);
    double* ox = x == 0 ? o[0] : (x == 1 ? o[1] : o[2]);
    const double* ux = x == 0 ? uu[0] : (x == 1 ? uu[1] : uu[2]);
    double out[Np];
    // volume: sum_r g[r][x] (D_r p)
#pragma unroll
    for (int i = 0; i < Np; ++i) {
      double a = 0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double d = 0;
#pragma unroll
        for (int m = 0; m < Np; ++m) d = fma(Cc[L.off_D + (r * Np + i) * Np + m], pp[m], d);
        a = fma(gx[r], d, a);
      }
      out[i] = a;
    }
    // faces: LIFT (fscale n_x (p* - p-)), interior only
#pragma unroll
    for (int f = 0; f < Nf; ++f) {
      const int nb = nbr[f * stride + e];
      if (nb < 0) continue;
      const int tbase = L.tab_nbr + (f * Nf * L.nperm + code[f * stride + e]) * Nfp;
      double c = 0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) c += face_alpha(DIM, f, r) * gx[r];
      for (int j = 0; j < Nfp; ++j) {
        const int pk = tb[tbase + j];
        const double jump = 0.5 * c * (p[(int64_t)(pk >> 6) * stride + nb] - pp[FM.v[f * Nfp + j]]);
#pragma unroll
        for (int i = 0; i < Np; ++i) out[i] = fma(Lc[i * NF + f * Nfp + j], jump, out[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < Np; ++i) {
      const int64_t k = i * stride + e;
      ox[k] = ux ? ux[k] - dt * out[i] : out[i];
    }
  }
}

}  // namespace

void launch_grad_central(int dim, int N, int64_t E, int64_t stride, const double* G, const int* nbr,
                         const int* code, const double* C64, const int* tab, const double* lift,
                         const double* p, double dt, const double* const u[3], double* const out[3],
                         hipStream_t s) {
  if (E == 0) return;
  const dim3 grid((unsigned)((E + 255) / 256)), block(256);
#define INSP_GRAD(D, NN)                                                                          \
  hipLaunchKernelGGL((k_grad_central<D, NN>), grid, block, 0, s, E, stride, G, nbr, code, C64,  \
                     tab, lift, p, dt, u[0], u[1], u[2], out[0], out[1], out[2])
  if (dim == 2) {
    switch (N) {
      case 1: INSP_GRAD(2, 1); break;
      case 2: INSP_GRAD(2, 2); break;
      case 3: INSP_GRAD(2, 3); break;
      case 4: INSP_GRAD(2, 4); break;
      default: throw std::runtime_error("unsupported order");
    }
  } else {
    switch (N) {
      case 1: INSP_GRAD(3, 1); break;
      case 2: INSP_GRAD(3, 2); break;
      case 3: INSP_GRAD(3, 3); break;
      case 4: INSP_GRAD(3, 4); break;
      default: throw std::runtime_error("unsupported order");
    }
  }
#undef INSP_GRAD
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess)
    throw std::runtime_error(std::string("k_grad_central launch failed: ") + hipGetErrorString(err));
}

}  // namespace insp
