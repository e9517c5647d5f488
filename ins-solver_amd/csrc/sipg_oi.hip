// Over-integrated variable-coefficient SIPG operator (the reference's
// FactorPoissonMatrixFreeDiagOI{2D,3D}, selected by [pressure-solve] over_int,
// src/solvers/2d/mp_ins_solver.cpp:98-100,580-600): the coefficient is given at the cubature
// points of the cells (kv) and of their faces (ks), and every integral of the operator is taken
// with those rules (oi.h). Element-centric like the nodal operator: a trace pass stores, for
// every face point of every cell, u and T = sJ k n.grad u (own outward normal); the element
// pass gathers the neighbour's values at the matching point (a permutation of the symmetric
// face rule), then
//
//   out_i = sum_q wv_q kv_q grad phi_i . Kt grad u (x_q)                    (volume, Kt = J grad xi grad xi^T)
//         + sum_f sum_q ws_q [ phi_i F + (beta_f . grad phi_i) sig ](x_fq)   (faces)
//   F = (tauS [u] - (T- - T+)) / 2,  sig = -ks [u] / 2,  beta_f = Kt alpha_f = sJ grad xi . n
//
// Dirichlet ghost u+ = -u-, T+ = -T- (vmf_2d_mult_avg_jump.h:32-38); Neumann faces none.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "dg_layout.h"
#include "kernels.h"
#include "reduce.h"
#include "sipg_cell.h"

namespace insp {

namespace {

// Points of the rules (oi.h): triangle volume / 3D faces 3, 6, 12, 16; edges 2N+1; tetrahedron
// volume (N+2)^3.
constexpr int oi_nv(int dim, int N) {
  return dim == 2 ? (N == 1 ? 3 : N == 2 ? 6 : N == 3 ? 12 : 16) : (N + 2) * (N + 2) * (N + 2);
}
constexpr int oi_ns(int dim, int N) {
  return dim == 2 ? 2 * N + 1 : (N == 1 ? 3 : N == 2 ? 6 : N == 3 ? 12 : 16);
}

template <int DIM, int N>
struct OiL {
  static constexpr int Np = DIM == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
  static constexpr int Nf = DIM + 1, nv = oi_nv(DIM, N), ns = oi_ns(DIM, N);
  static constexpr int off_GV = 0, off_wv = off_GV + DIM * nv * Np, off_IS = off_wv + nv;
  static constexpr int off_GS = off_IS + Nf * ns * Np, off_ws = off_GS + Nf * DIM * ns * Np;
  static constexpr int size = off_ws + ns;
  static constexpr int nperm = DIM == 2 ? 2 : 6;
};

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}

template <int DIM, typename T>
__device__ __forceinline__ void face_beta(const T* K, int f, T* beta) {
#pragma unroll
  for (int s = 0; s < DIM; ++s) {
    T b = 0;
#pragma unroll
    for (int r = 0; r < DIM; ++r) b += (T)face_alpha(DIM, f, r) * K[kidx(DIM, r, s)];
    beta[s] = b;
  }
}

template <int DIM, int N, typename T>
__global__ void __launch_bounds__(256)
k_trace_oi(int64_t E, int64_t stride, const T* __restrict__ u, const double* __restrict__ Kt,
           const T* __restrict__ ks, const T* __restrict__ Cp, T* __restrict__ tU, T* __restrict__ tT) {
  using L = OiL<DIM, N>;
  const cptr<T> C = as_const(Cp);
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  constexpr int NK = DIM == 2 ? 3 : 6;
  T uu[L::Np], K[NK];
#pragma unroll
  for (int m = 0; m < L::Np; ++m) uu[m] = u[m * stride + e];
#pragma unroll
  for (int k = 0; k < NK; ++k) K[k] = (T)Kt[k * stride + e];
#pragma unroll
  for (int f = 0; f < L::Nf; ++f) {
    T beta[DIM];
    face_beta<DIM>(K, f, beta);
    for (int q = 0; q < L::ns; ++q) {
      T U = 0, Tn = 0;
#pragma unroll
      for (int m = 0; m < L::Np; ++m) U = fma(C[L::off_IS + (f * L::ns + q) * L::Np + m], uu[m], U);
#pragma unroll
      for (int s = 0; s < DIM; ++s) {
        T g = 0;
#pragma unroll
        for (int m = 0; m < L::Np; ++m) g = fma(C[L::off_GS + ((f * DIM + s) * L::ns + q) * L::Np + m], uu[m], g);
        Tn = fma(beta[s], g, Tn);
      }
      const int64_t idx = (int64_t)(f * L::ns + q) * stride + e;
      tU[idx] = U;
      tT[idx] = ks[idx] * Tn;
    }
  }
}

template <int DIM, int N, typename T>
__global__ void __launch_bounds__(256)
k_apply_oi(int64_t E, int64_t stride, const T* __restrict__ u, const T* __restrict__ tU,
           const T* __restrict__ tT, const double* __restrict__ Kt, const double* __restrict__ tauS,
           const int* __restrict__ nbr, const int* __restrict__ code, const int* __restrict__ bctype,
           const T* __restrict__ kv, const T* __restrict__ ks, const T* __restrict__ Cp,
           const int* __restrict__ qmapp, T* __restrict__ out, double* __restrict__ partial) {
  using L = OiL<DIM, N>;
  const cptr<T> C = as_const(Cp);
  const cptr<int> qmap = as_const(qmapp);
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (e < E) {
    constexpr int NK = DIM == 2 ? 3 : 6;
    T uu[L::Np], acc[L::Np], K[NK];
#pragma unroll
    for (int m = 0; m < L::Np; ++m) {
      uu[m] = u[m * stride + e];
      acc[m] = 0;
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) K[k] = (T)Kt[k * stride + e];
    // volume: k at the cubature points
    for (int q = 0; q < L::nv; ++q) {
      T g[DIM];
#pragma unroll
      for (int s = 0; s < DIM; ++s) {
        T a = 0;
#pragma unroll
        for (int m = 0; m < L::Np; ++m) a = fma(C[L::off_GV + (s * L::nv + q) * L::Np + m], uu[m], a);
        g[s] = a;
      }
      const T wk = C[L::off_wv + q] * kv[(int64_t)q * stride + e];
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        T h = 0;
#pragma unroll
        for (int s = 0; s < DIM; ++s) h = fma(K[kidx(DIM, r, s)], g[s], h);
        h *= wk;
#pragma unroll
        for (int i = 0; i < L::Np; ++i) acc[i] = fma(C[L::off_GV + (r * L::nv + q) * L::Np + i], h, acc[i]);
      }
    }
    // faces at the face points
#pragma unroll
    for (int f = 0; f < L::Nf; ++f) {
      const int nb = nbr[f * stride + e];
      int mode = 0, fn = 0, p = 0;
      if (nb >= 0) {
        const int c = code[f * stride + e];
        fn = c / L::nperm;
        p = c - fn * L::nperm;
      } else {
        mode = bctype[-nb - 1] == 0 ? 1 : 2;
      }
      if (mode == 2) continue;
      T beta[DIM];
      face_beta<DIM>(K, f, beta);
      const T ts = (T)tauS[f * stride + e];
      for (int q = 0; q < L::ns; ++q) {
        const int64_t own = (int64_t)(f * L::ns + q) * stride + e;
        const T Um = tU[own], Tm = tT[own];
        T Up, Tp;
        if (mode == 0) {
          const int64_t nidx = (int64_t)(fn * L::ns + qmap[p * L::ns + q]) * stride + nb;
          Up = tU[nidx];
          Tp = tT[nidx];
        } else {
          Up = -Um;
          Tp = -Tm;
        }
        const T w = C[L::off_ws + q];
        const T jump = Um - Up;
        const T F = w * (T)0.5 * (ts * jump - (Tm - Tp));
        const T sig = w * (T)-0.5 * ks[own] * jump;
#pragma unroll
        for (int i = 0; i < L::Np; ++i) {
          T dn = 0;
#pragma unroll
          for (int s = 0; s < DIM; ++s) dn = fma(beta[s], C[L::off_GS + ((f * DIM + s) * L::ns + q) * L::Np + i], dn);
          acc[i] = fma(C[L::off_IS + (f * L::ns + q) * L::Np + i], F, fma(dn, sig, acc[i]));
        }
      }
    }
#pragma unroll
    for (int i = 0; i < L::Np; ++i) {
      out[i * stride + e] = acc[i];
      dotv += (double)uu[i] * (double)acc[i];
    }
  }
  if (partial) {
    const double t = block_sum1<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// Exact diagonal of the over-integrated operator.
template <int DIM, int N>
__global__ void __launch_bounds__(256)
k_diag_oi(int64_t E, int64_t stride, const double* __restrict__ Kt, const double* __restrict__ tauS,
          const int* __restrict__ nbr, const int* __restrict__ bctype, const double* __restrict__ kv,
          const double* __restrict__ ks, const double* __restrict__ C, double* __restrict__ diag) {
  using L = OiL<DIM, N>;
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  constexpr int NK = DIM == 2 ? 3 : 6;
  double K[NK];
  for (int k = 0; k < NK; ++k) K[k] = Kt[k * stride + e];
  for (int i = 0; i < L::Np; ++i) {
    double d = 0;
    for (int q = 0; q < L::nv; ++q) {
      double g[DIM];
      for (int s = 0; s < DIM; ++s) g[s] = C[L::off_GV + (s * L::nv + q) * L::Np + i];
      double a = 0;
      for (int r = 0; r < DIM; ++r)
        for (int s = 0; s < DIM; ++s) a += g[r] * K[kidx(DIM, r, s)] * g[s];
      d += C[L::off_wv + q] * kv[(int64_t)q * stride + e] * a;
    }
    for (int f = 0; f < L::Nf; ++f) {
      const int nb = nbr[f * stride + e];
      double scale = 0.5;
      if (nb < 0) {
        if (bctype[-nb - 1] != 0) continue;
        scale = 1.0;
      }
      double beta[DIM];
      face_beta<DIM>(K, f, beta);
      const double ts = tauS[f * stride + e];
      for (int q = 0; q < L::ns; ++q) {
        const double phi = C[L::off_IS + (f * L::ns + q) * L::Np + i];
        double dn = 0;
        for (int s = 0; s < DIM; ++s) dn += beta[s] * C[L::off_GS + ((f * DIM + s) * L::ns + q) * L::Np + i];
        d += scale * C[L::off_ws + q] *
             (ts * phi * phi - 2.0 * ks[(int64_t)(f * L::ns + q) * stride + e] * phi * dn);
      }
    }
    diag[i * stride + e] = d;
  }
}

dim3 grid_oi(int64_t E) {
  const int64_t nb = (E + 255) / 256;
  return dim3((unsigned)nb);
}

}  // namespace

#define INSP_OI_DISPATCH(dim, N, CALL)                                  \
  do {                                                                  \
    if ((dim) == 2) {                                                   \
      switch (N) {                                                      \
        case 1: { constexpr int D_ = 2, N_ = 1; CALL; } break;          \
        case 2: { constexpr int D_ = 2, N_ = 2; CALL; } break;          \
        case 3: { constexpr int D_ = 2, N_ = 3; CALL; } break;          \
        case 4: { constexpr int D_ = 2, N_ = 4; CALL; } break;          \
        default: throw std::runtime_error("unsupported order");         \
      }                                                                 \
    } else {                                                            \
      switch (N) {                                                      \
        case 1: { constexpr int D_ = 3, N_ = 1; CALL; } break;          \
        case 2: { constexpr int D_ = 3, N_ = 2; CALL; } break;          \
        case 3: { constexpr int D_ = 3, N_ = 3; CALL; } break;          \
        case 4: { constexpr int D_ = 3, N_ = 4; CALL; } break;          \
        default: throw std::runtime_error("unsupported order");         \
      }                                                                 \
    }                                                                   \
  } while (0)

bool oi_shape_matches(int dim, int N, int nv, int ns, int table_size) {
  bool ok = false;
  INSP_OI_DISPATCH(dim, N, ok = (OiL<D_, N_>::nv == nv && OiL<D_, N_>::ns == ns && OiL<D_, N_>::size == table_size));
  return ok;
}

template <typename T>
void launch_apply_oi(const MeshView& m, const LevelView& l, const OiView<T>& o, const T* u, T* out,
                     hipStream_t s, double* partial) {
  if (m.E == 0) return;
  INSP_OI_DISPATCH(m.dim, l.N,
                   hipLaunchKernelGGL((k_trace_oi<D_, N_, T>), grid_oi(m.E), dim3(256), 0, s, m.E, m.stride, u,
                                      m.Kt, o.ks, o.C, o.tU, o.tT));
  check("k_trace_oi");
  INSP_OI_DISPATCH(m.dim, l.N,
                   hipLaunchKernelGGL((k_apply_oi<D_, N_, T>), grid_oi(m.E), dim3(256), 0, s, m.E, m.stride, u,
                                      o.tU, o.tT, m.Kt, l.tauS, m.nbr, m.code, m.bctype, o.kv, o.ks, o.C,
                                      o.qmap, out, partial));
  check("k_apply_oi");
}

void launch_diag_oi(const MeshView& m, const LevelView& l, const OiView<double>& o, double* diag,
                    hipStream_t s) {
  if (m.E == 0) return;
  INSP_OI_DISPATCH(m.dim, l.N,
                   hipLaunchKernelGGL((k_diag_oi<D_, N_>), grid_oi(m.E), dim3(256), 0, s, m.E, m.stride, m.Kt,
                                      l.tauS, m.nbr, m.bctype, o.kv, o.ks, o.C, diag));
  check("k_diag_oi");
}

template void launch_apply_oi<double>(const MeshView&, const LevelView&, const OiView<double>&, const double*,
                                      double*, hipStream_t, double*);
template void launch_apply_oi<float>(const MeshView&, const LevelView&, const OiView<float>&, const float*,
                                     float*, hipStream_t, double*);

}  // namespace insp
