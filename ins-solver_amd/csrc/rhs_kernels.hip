// Caller-side pieces of the pressure solve that the reference runs around
// LinearSolver::solve (src/solvers/2d/ins_solver.cpp:310-319):
//   ins_pressure_rhs_2d.h:1-12  div <- -div/dt,  dPdNOld <- b0 dPdN + b1 dPdNOld
//   op2_gemv(LIFT, dPdNOld) added to div, then mesh->mass
// fused into one pass per cell: rhs = J M (-div/dt) + J EMAT g, g = b0 dPdN + b1 dPdNOld,
// since J M (M^-1 EMAT g) = J EMAT g. Face data are cells x (Nf*Nfp) in the owner's
// face-mask order, SoA with the solver's stride.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.h"

namespace insp {

namespace {

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("rhs kernel launch failed: ") + what + ": " +
                             hipGetErrorString(e));
}

constexpr int kB = 256;

template <int NP, int NFACE>
__global__ void __launch_bounds__(kB)
k_pressure_rhs(int64_t E, int64_t stride, const double* __restrict__ J, const double* __restrict__ M,
               const double* __restrict__ EMAT, double b0, double b1, double dt,
               const double* __restrict__ dPdN, double* __restrict__ dPdNOld,
               double* __restrict__ div) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= E) return;
  const double inv_dt = 1.0 / dt;
  double g[NFACE], v[NP];
#pragma unroll
  for (int q = 0; q < NFACE; ++q) {
    const int64_t k = q * stride + e;
    g[q] = b0 * dPdN[k] + b1 * dPdNOld[k];
    dPdNOld[k] = g[q];
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) v[j] = -div[j * stride + e] * inv_dt;
  const double jac = J[e];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NP; ++j) s = fma(M[i * NP + j], v[j], s);
#pragma unroll
    for (int q = 0; q < NFACE; ++q) s = fma(EMAT[i * NFACE + q], g[q], s);
    div[i * stride + e] = jac * s;
  }
}

}  // namespace

void launch_pressure_rhs(int dim, int N, int64_t E, int64_t stride, const double* J, const double* M,
                         const double* EMAT, double b0, double b1, double dt, const double* dPdN,
                         double* dPdNOld, double* div, hipStream_t s) {
  if (E == 0) return;
  const dim3 g((unsigned)((E + kB - 1) / kB)), b(kB);
#define INSP_RHS(NP, NF) \
  hipLaunchKernelGGL((k_pressure_rhs<NP, NF>), g, b, 0, s, E, stride, J, M, EMAT, b0, b1, dt, dPdN, dPdNOld, div)
  if (dim == 2) {
    switch (N) {
      case 1: INSP_RHS(3, 6); break;
      case 2: INSP_RHS(6, 9); break;
      case 3: INSP_RHS(10, 12); break;
      case 4: INSP_RHS(15, 15); break;
      default: throw std::runtime_error("unsupported order");
    }
  } else {
    switch (N) {
      case 1: INSP_RHS(4, 12); break;
      case 2: INSP_RHS(10, 24); break;
      case 3: INSP_RHS(20, 40); break;
      case 4: INSP_RHS(35, 60); break;
      default: throw std::runtime_error("unsupported order");
    }
  }
#undef INSP_RHS
  check("k_pressure_rhs");
}

}  // namespace insp
