// Mass term of the viscous (Helmholtz) operator and the inverse-mass preconditioner.
//
//   out += J M (mm o u)     per-node factor  (fvmf_2d_mult_mm_geof.h:1-26, 3D twin)
//   out += mm J M u         constant factor  (vmf_2d_mult_mm_geof.h:1-26, 3D twin)
//   diag += mm_i J M_ii     (fvmf_2d_mm_diag.h:1-9)
//   z = s (J M)^-1 r  or  z_i = ((J M)^-1 r)_i / f_i
//                           (viscous_solver.cpp:334-366, cg_recp_factor_inv_mass.h)
//
// One thread per cell; the Np x Np matrices are wave-uniform and read through the
// scalar cache. The mass pass runs after the element pass of the SIPG operator on the
// same 256-cell blocks, so its share of u.(A u) is added to the same partial slots.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "dg_layout.h"
#include "kernels.h"
#include "reduce.h"
#include "sipg_cell.h"

namespace insp {

namespace {

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("mass kernel launch failed: ") + what + ": " +
                             hipGetErrorString(e));
}

template <int NP, typename T>
__global__ void __launch_bounds__(256)
k_mass_add(int64_t E, int64_t stride, const double* __restrict__ J, const T* __restrict__ M,
           const T* __restrict__ mm, T mmc, const T* __restrict__ u, T* __restrict__ out,
           double* __restrict__ partial) {
  const cptr<T> Mc = as_const(M);
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (e < E) {
    T v[NP], uu[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      uu[j] = u[j * stride + e];
      v[j] = (mm ? mm[j * stride + e] : mmc) * uu[j];
    }
    const T Je = (T)J[e];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      T a = 0;
#pragma unroll
      for (int j = 0; j < NP; ++j) a = fma(Mc[i * NP + j], v[j], a);
      a *= Je;
      out[i * stride + e] += a;
      dotv += (double)uu[i] * (double)a;
    }
  }
  if (partial) {
    const double t = block_sum1<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] += t;
  }
}

template <int NP>
__global__ void __launch_bounds__(256)
k_mass_diag(int64_t E, int64_t stride, const double* __restrict__ J, const double* __restrict__ M,
            const double* __restrict__ mm, double mmc, double* __restrict__ diag) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
#pragma unroll
  for (int i = 0; i < NP; ++i)
    diag[i * stride + e] += (mm ? mm[i * stride + e] : mmc) * J[e] * M[i * NP + i];
}

template <int NP>
__global__ void __launch_bounds__(256)
k_inv_mass(int64_t E, int64_t stride, const double* __restrict__ J, const double* __restrict__ Minv,
           double s, const double* __restrict__ recp, const double* __restrict__ r,
           double* __restrict__ z) {
  const cptr<double> Mi = as_const(Minv);
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double rr[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) rr[j] = r[j * stride + e];
  const double inv_j = 1.0 / J[e];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double a = 0;
#pragma unroll
    for (int j = 0; j < NP; ++j) a = fma(Mi[i * NP + j], rr[j], a);
    a *= inv_j;
    z[i * stride + e] = recp ? a / recp[i * stride + e] : a * s;
  }
}

inline dim3 grid_of(int64_t E) { return dim3((unsigned)((E + 255) / 256)); }

#define INSP_NP_SWITCH(np, CALL)                                 \
  switch (np) {                                                  \
    case 3: { constexpr int NP_ = 3; CALL; } break;              \
    case 4: { constexpr int NP_ = 4; CALL; } break;              \
    case 6: { constexpr int NP_ = 6; CALL; } break;              \
    case 10: { constexpr int NP_ = 10; CALL; } break;            \
    case 15: { constexpr int NP_ = 15; CALL; } break;            \
    case 20: { constexpr int NP_ = 20; CALL; } break;            \
    case 35: { constexpr int NP_ = 35; CALL; } break;            \
    default: throw std::runtime_error("unsupported number of nodes"); \
  }

}  // namespace

template <typename T>
void launch_mass_add(const MeshView& m, const LevelView& l, int np, const T* mm, T mmc, const T* u,
                     T* out, hipStream_t s, double* partial) {
  if (m.E == 0) return;
  const OpLayout L = op_layout(m.dim, l.N);
  const T* M = static_cast<const T*>(l.C) + L.off_M;
  INSP_NP_SWITCH(np, hipLaunchKernelGGL((k_mass_add<NP_, T>), grid_of(m.E), dim3(256), 0, s, m.E,
                                        m.stride, m.J, M, mm, mmc, u, out, partial));
  check("k_mass_add");
}

void launch_mass_diag(const MeshView& m, const LevelView& l, int np, const double* mm, double mmc,
                      double* diag, hipStream_t s) {
  if (m.E == 0) return;
  const OpLayout L = op_layout(m.dim, l.N);
  const double* M = l.C64 + L.off_M;
  INSP_NP_SWITCH(np, hipLaunchKernelGGL((k_mass_diag<NP_>), grid_of(m.E), dim3(256), 0, s, m.E,
                                        m.stride, m.J, M, mm, mmc, diag));
  check("k_mass_diag");
}

void launch_inv_mass(int np, int64_t E, int64_t stride, const double* J, const double* Minv,
                     double scale, const double* recp, const double* r, double* z, hipStream_t s) {
  if (E == 0) return;
  INSP_NP_SWITCH(np, hipLaunchKernelGGL((k_inv_mass<NP_>), grid_of(E), dim3(256), 0, s, E, stride,
                                        J, Minv, scale, recp, r, z));
  check("k_inv_mass");
}

template void launch_mass_add<double>(const MeshView&, const LevelView&, int, const double*, double,
                                      const double*, double*, hipStream_t, double*);
template void launch_mass_add<float>(const MeshView&, const LevelView&, int, const float*, float,
                                     const float*, float*, hipStream_t, double*);

}  // namespace insp
