// Matrix-free SIPG Poisson operator on gfx950.
//
// The operator is the reference's slip-matrix viscous operator with mass
// factor 0 and one component (src/solvers/2d/slip_matrix/factor_viscous_matrix.cpp:87-209,
// 3D twin src/solvers/3d/slip_matrix/factor_viscous_matrix.cpp), rewritten
// element-centrically so that no face colouring and no atomics are needed:
//
//   out = sum_r D_r^T M (k * sum_s Kt_rs D_s u)                     (volume)
//       + sum_f EMAT_f F_f + sum_f sum_r beta_fr (D_r^T EMAT_f) sig_f  (faces)
//
//   F_f   = sJ (tau * [u]/2 - n.{k grad u})     (fvmf_2d_mult_flux.h:7-10)
//   sig_f = -sJ k [u]/2 (times n, lifted)        (fvmf_2d_mult_flux.h:11-14)
//
// with Kt = J * grad(xi) grad(xi)^T, beta_f = Kt alpha_f = sJ * grad(xi).n_f,
// [u] = u- - u+, and the Dirichlet ghost u+ = -u-, (k grad u)+ = (k grad u)-
// (vmf_2d_mult_avg_jump.h:32-38). Neumann faces contribute nothing.
//
// Two passes: the trace pass stores, for every cell face node, the sJ-scaled
// normal flux trace T = sJ n.(k grad u) (a cell's own outward normal); the
// element pass gathers the neighbour's T and u by index and permutation.
// One thread owns one cell: reference matrices are wave-uniform and are read
// through the scalar cache; the cell's nodal values stay in VGPRs.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "dg_layout.h"
#include "kernels.h"
#include "reduce.h"
#include "refelem.h"
#include "sipg_cell.h"

namespace insp {

namespace {
inline void launch_check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}
}  // namespace

namespace {

template <int DIM, int N, typename T, bool VAR>
__global__ void __launch_bounds__(256)
k_trace(int64_t E, int64_t stride, const T* __restrict__ u, const double* __restrict__ Kt,
        const T* __restrict__ kfac, const T* __restrict__ C, T* __restrict__ tr) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e < E) trace_cell<DIM, N, T, VAR>(e, stride, u, Kt, kfac, as_const(C), tr);
}

template <int DIM, int N, typename T, bool VAR, bool PACK>
__global__ void __launch_bounds__(256)
k_apply(int64_t E, int64_t stride, const T* __restrict__ u, const T* __restrict__ tr,
        const double* __restrict__ Kt, const double* __restrict__ tauS,
        const int* __restrict__ nbr, const int* __restrict__ code,
        const int* __restrict__ bctype, const T* __restrict__ kfac, const T* __restrict__ C,
        const int* __restrict__ tab, T* __restrict__ out, double* __restrict__ partial) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (e < E)
    dotv = apply_cell<DIM, N, T, VAR, PACK>(e, stride, u, tr, Kt, tauS, nbr, code, bctype, kfac, as_const(C),
                                      as_const(tab), out);
  // optional fused u.(A u) partial sum for CG step lengths
  if (partial) {
    const double t = block_sum1<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}


template <int DIM, typename T, bool VAR, bool PACK, bool KAOS, bool VCM = false>
__global__ void __launch_bounds__(256)
k_apply_p1(int64_t E, int64_t stride, const T* __restrict__ u, const double* __restrict__ Kt,
           const double* __restrict__ tauS, const int* __restrict__ nbr,
           const int* __restrict__ code, const int* __restrict__ bctype,
           const T* __restrict__ kfac, const T* __restrict__ C, const int* __restrict__ tab,
           T* __restrict__ out, double* __restrict__ partial, P1NodeBits bits) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (e < E)
    dotv = apply_p1_cell<DIM, T, VAR, PACK, T*, NoSlip, BitNodes, KAOS, VCM>(
        e, stride, u, Kt, tauS, nbr, code, bctype, kfac, as_const(C), as_const(tab), out, NoSlip{},
        BitNodes{bits});
  if (partial) {
    const double t = block_sum1<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// Element pass with the Chebyshev step of the smoother fused into its output (orders >= 2).
template <int DIM, int N, typename T, bool VAR, bool PACK>
__global__ void __launch_bounds__(256)
k_apply_cheb(int64_t E, int64_t stride, const T* __restrict__ u, const T* __restrict__ tr,
             const double* __restrict__ Kt, const double* __restrict__ tauS,
             const int* __restrict__ nbr, const int* __restrict__ code,
             const int* __restrict__ bctype, const T* __restrict__ kfac, const T* __restrict__ C,
             const int* __restrict__ tab, const T* __restrict__ f, const T* __restrict__ inv,
             const T* prev, T* x_next, T c1, T c2) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e < E)
    (void)apply_cell<DIM, N, T, VAR, PACK>(e, stride, u, tr, Kt, tauS, nbr, code, bctype, kfac,
                                           as_const(C), as_const(tab),
                                           ChebOut<T>{f, inv, prev, x_next, c1, c2});
}

// Order-1 element pass with the Chebyshev step fused into its output.
template <int DIM, typename T, bool VAR, bool PACK, bool KAOS>
__global__ void __launch_bounds__(256)
k_apply_p1_cheb(int64_t E, int64_t stride, const T* __restrict__ u, const double* __restrict__ Kt,
                const double* __restrict__ tauS, const int* __restrict__ nbr,
                const int* __restrict__ code, const int* __restrict__ bctype,
                const T* __restrict__ kfac, const T* __restrict__ C, const int* __restrict__ tab,
                const T* __restrict__ f, const T* __restrict__ inv, const T* prev, T* x_next,
                T c1, T c2, P1NodeBits bits) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e < E)
    (void)apply_p1_cell<DIM, T, VAR, PACK, ChebOut<T>, NoSlip, BitNodes, KAOS>(
        e, stride, u, Kt, tauS, nbr, code, bctype, kfac, as_const(C), as_const(tab),
        ChebOut<T>{f, inv, prev, x_next, c1, c2}, NoSlip{}, BitNodes{bits});
}

// Element pass of one component of the vector (velocity) operator with slip faces (FP64;
// orders >= 2 read the face traces, order 1 recomputes them). kfac may be null (k = 1).
template <int DIM, int N>
__global__ void __launch_bounds__(256)
k_apply_slip(int64_t E, int64_t stride, const double* __restrict__ u, const double* __restrict__ tr,
             const double* __restrict__ Kt, const double* __restrict__ tauS,
             const int* __restrict__ nbr, const int* __restrict__ code,
             const int* __restrict__ bctype, const double* __restrict__ kfac,
             const double* __restrict__ C, const int* __restrict__ tab, double* __restrict__ out,
             double* __restrict__ partial, SlipFaces<double> slip) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (e < E) {
    if constexpr (N == 1)
      dotv = apply_p1_cell<DIM, double, 2, false, double*, SlipFaces<double>>(
          e, stride, u, Kt, tauS, nbr, code, bctype, kfac, as_const(C), as_const(tab), out, slip);
    else
      dotv = apply_cell<DIM, N, double, 2, false, double*, SlipFaces<double>>(
          e, stride, u, tr, Kt, tauS, nbr, code, bctype, kfac, as_const(C), as_const(tab), out, slip);
  }
  if (partial) {
    const double t = block_sum1<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

template <int DIM, int N, bool VAR>
__global__ void __launch_bounds__(256)
k_diag(int64_t E, int64_t stride, const double* __restrict__ Kt, const double* __restrict__ tauS,
       const int* __restrict__ nbr, const int* __restrict__ bctype,
       const double* __restrict__ kfac, const double* __restrict__ C,
       const int* __restrict__ tab, double* __restrict__ diag,
       const double* __restrict__ bn, int comp) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf;
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double K[L.NK];
#pragma unroll
  for (int k = 0; k < L.NK; ++k) K[k] = Kt[k * stride + e];
  double beta[Nf][DIM], ts[Nf], wf[Nf];
  int mode[Nf];
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
#pragma unroll
    for (int s = 0; s < DIM; ++s) {
      double b = 0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) b += face_alpha(DIM, f, r) * K[kidx(DIM, r, s)];
      beta[f][s] = b;
    }
    const int nb = nbr[f * stride + e];
    // slip faces of the vector operator (bn given): the Dirichlet terms of component comp,
    // scaled by n_comp^2 (the part of the projected jump and trace that is its own)
    wf[f] = 1.0;
    if (nb >= 0) {
      mode[f] = 0;
    } else {
      const int bt = bctype[-nb - 1];
      mode[f] = bt == 0 ? 1 : 2;
      if (bn && bt == 2) {
        const double nc = bn[(int64_t)(-nb - 1) * DIM + comp];
        mode[f] = 1;
        wf[f] = nc * nc;
      }
    }
    ts[f] = tauS[f * stride + e];
  }
  for (int i = 0; i < Np; ++i) {
    double d = 0;
    for (int m = 0; m < Np; ++m) {
      double dm[DIM];
#pragma unroll
      for (int s = 0; s < DIM; ++s) dm[s] = C[L.off_D + (s * Np + m) * Np + i];
      double a = 0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double h = 0;
#pragma unroll
        for (int s = 0; s < DIM; ++s) h += K[kidx(DIM, r, s)] * dm[s];
        a += C[L.off_Bt + (r * Np + m) * Np + i] * h;
      }
      d += (VAR ? kfac[m * stride + e] : 1.0) * a;
    }
#pragma unroll
    for (int f = 0; f < Nf; ++f) {
      if (mode[f] == 2) continue;
      const double scale = mode[f] == 0 ? 0.5 : 1.0;
      const int li = tab[L.tab_finv + f * Np + i];
      for (int j = 0; j < Nfp; ++j) {
        const int p = tab[L.tab_fmask + f * Nfp + j];
        const double kp = VAR ? kfac[p * stride + e] : 1.0;
        const double um = p == i ? 1.0 : 0.0;
        double tm = 0;
#pragma unroll
        for (int s = 0; s < DIM; ++s) tm += beta[f][s] * C[L.off_D + (s * Np + p) * Np + i];
        tm *= kp;
        const double F = wf[f] * scale * (ts[f] * um - tm);
        const double sig = -wf[f] * scale * kp * um;
        if (li >= 0) d += C[L.off_MfT + (f * Nfp + j) * Nfp + li] * F;
#pragma unroll
        for (int r = 0; r < DIM; ++r)
          d += C[L.off_Wt + ((f * DIM + r) * Nfp + j) * Np + i] * beta[f][r] * sig;
      }
    }
    diag[i * stride + e] = d;
  }
}

template <int DIM, int N, bool VAR>
__global__ void __launch_bounds__(256)
k_tau(int64_t E, int64_t stride, const double* __restrict__ Kt, const double* __restrict__ J,
      const int* __restrict__ nbr, const int* __restrict__ code,
      const double* __restrict__ kfac, const int* __restrict__ tab, int pen_order,
      double* __restrict__ tauS) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Nfp = L.Nfp, Nf = L.Nf;
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  auto fscale = [&](int64_t c, int f) {
    double q = 0;
    for (int r = 0; r < DIM; ++r)
      for (int s = 0; s < DIM; ++s)
        q += face_alpha(DIM, f, r) * face_alpha(DIM, f, s) * Kt[kidx(DIM, r, s) * stride + c];
    return sqrt(q / J[c]);
  };
  auto kmax = [&](int64_t c, int f) {
    if (!VAR) return 1.0;
    double k = -1e300;
    for (int j = 0; j < Nfp; ++j) k = fmax(k, kfac[tab[L.tab_fmask + f * Nfp + j] * stride + c]);
    return k;
  };
  const double pen = 2.0 * (pen_order + 1) * (pen_order + 2);
  for (int f = 0; f < Nf; ++f) {
    const double fs = fscale(e, f);
    double t = fs * kmax(e, f);
    const int nb = nbr[f * stride + e];
    if (nb >= 0) {
      const int fn = code[f * stride + e] / L.nperm;
      t = fmax(t, fscale(nb, fn) * kmax(nb, fn));
    }
    tauS[f * stride + e] = pen * t * fs * J[e];
  }
}

template <int DIM, int N, bool VAR>
__global__ void __launch_bounds__(256)
k_apply_bc(int64_t E, int64_t stride, const double* __restrict__ Kt, const double* __restrict__ J,
           const double* __restrict__ tauS, const int* __restrict__ nbr,
           const int* __restrict__ bctype, const double* __restrict__ bc,
           const double* __restrict__ kfac, const double* __restrict__ C,
           const int* __restrict__ tab, double* __restrict__ rhs) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf;
  constexpr FmaskTable FM = make_fmask(DIM, N);
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  bool any = false;
  for (int f = 0; f < Nf; ++f) any |= nbr[f * stride + e] < 0;
  if (!any) return;
  double acc[Np];
#pragma unroll
  for (int i = 0; i < Np; ++i) acc[i] = rhs[i * stride + e];
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
    const int nb = nbr[f * stride + e];
    if (nb >= 0) continue;
    const int bf = -nb - 1;
    const int type = bctype[bf];
    if (type != 0 && type != 1) continue;
    double t[Nfp];
#pragma unroll
    for (int l = 0; l < Nfp; ++l) {
      double a = 0;
      for (int j = 0; j < Nfp; ++j) a += C[L.off_MfT + (f * Nfp + j) * Nfp + l] * bc[(int64_t)bf * Nfp + j];
      t[l] = a;
    }
    if (type == 0) {
      // Dirichlet: tau sJ MM_F g - (n.grad)^T (k sJ MM_F g)   (fvmf_2d_apply_bc.h:24-67)
      double beta[DIM];
      for (int s = 0; s < DIM; ++s) {
        double b = 0;
        for (int r = 0; r < DIM; ++r) b += face_alpha(DIM, f, r) * Kt[kidx(DIM, r, s) * stride + e];
        beta[s] = b;
      }
      const double ts = tauS[f * stride + e];
#pragma unroll
      for (int l = 0; l < Nfp; ++l) {
        const int p = FM.v[f * Nfp + l];
        acc[p] += ts * t[l];
        const double kt = (VAR ? kfac[p * stride + e] : 1.0) * t[l];
#pragma unroll
        for (int i = 0; i < Np; ++i) {
          double dn = 0;
#pragma unroll
          for (int r = 0; r < DIM; ++r) dn += beta[r] * C[L.off_D + (r * Np + p) * Np + i];
          acc[i] -= dn * kt;
        }
      }
    } else {
      // Neumann: sJ MM_F g   (fvmf_2d_apply_bc.h:68-76)
      double q = 0;
      for (int r = 0; r < DIM; ++r)
        for (int s = 0; s < DIM; ++s)
          q += face_alpha(DIM, f, r) * face_alpha(DIM, f, s) * Kt[kidx(DIM, r, s) * stride + e];
      const double sJ = sqrt(q * J[e]);
#pragma unroll
      for (int l = 0; l < Nfp; ++l) acc[FM.v[f * Nfp + l]] += sJ * t[l];
    }
  }
#pragma unroll
  for (int i = 0; i < Np; ++i) rhs[i * stride + e] = acc[i];
}

template <int DIM, int N>
__global__ void __launch_bounds__(256)
k_mass(int64_t E, int64_t stride, const double* __restrict__ J, const double* __restrict__ C,
       const double* __restrict__ in, double* __restrict__ out) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Np = L.Np;
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double x[Np];
#pragma unroll
  for (int j = 0; j < Np; ++j) x[j] = in[j * stride + e];
  const double Je = J[e];
#pragma unroll
  for (int i = 0; i < Np; ++i) {
    double a = 0;
#pragma unroll
    for (int j = 0; j < Np; ++j) a = fma(C[L.off_M + i * Np + j], x[j], a);
    out[i * stride + e] = Je * a;
  }
}

template <typename T, int R, int K, bool ADD>
__global__ void __launch_bounds__(256)
k_elem_matvec(int64_t E, int64_t stride, const T* __restrict__ A, const T* __restrict__ x,
              T* __restrict__ y) {
  const int64_t e = xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  T xv[K];
#pragma unroll
  for (int j = 0; j < K; ++j) xv[j] = x[j * stride + e];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    T a = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) a = fma(A[i * K + j], xv[j], a);
    if (ADD) a += y[i * stride + e];
    y[i * stride + e] = a;
  }
}

inline dim3 grid_for(int64_t E) { return dim3((unsigned)((E + 255) / 256)); }

#define INSP_DISPATCH_N(dim, N, CALL)                                   \
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
    launch_check("SIPG kernel");                                        \
  } while (0)

}  // namespace

template <typename T>
void launch_trace(const MeshView& m, const LevelView& l, const T* u, const T* kfac, T* tr,
                  hipStream_t s) {
  if (m.E == 0 || l.N == 1) return;  // order 1 runs in one pass (k_apply_p1)
  const T* C = static_cast<const T*>(l.C);
  if (kfac)
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_trace<D_, N_, T, true>), grid_for(m.E), dim3(256), 0, s,
                                       m.E, m.stride, u, m.Kt, kfac, C, tr));
  else
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_trace<D_, N_, T, false>), grid_for(m.E), dim3(256), 0, s,
                                       m.E, m.stride, u, m.Kt, kfac, C, tr));
}

template <typename T>
void launch_apply(const MeshView& m, const LevelView& l, const T* u, const T* tr, const T* kfac,
                  T* out, hipStream_t s, double* partial) {
  if (m.E == 0) return;
  if (l.frag && mfma_apply_supported(m.dim, l.N)) {
    launch_apply_mfma<T>(m, l, u, tr, kfac, out, s, partial);
    return;
  }
  const T* C = static_cast<const T*>(l.C);
  if (l.N == 1) {
    // packed neighbour words when the mesh has them (one 4-byte load per face instead of two)
    const bool pack = m.nbc != nullptr;
    const int* nw = pack ? m.nbc : m.nbr;
    const P1NodeBits& bits = p1_node_bits(m.dim);
#define INSP_P1(D, V, P, A)                                                                       \
  hipLaunchKernelGGL((k_apply_p1<D, T, V, P, A>), grid_for(m.E), dim3(256), 0, s, m.E, m.stride, \
                     u, A ? m.Ka : m.Kt, l.tauS, nw, m.code, m.bctype, kfac, C, l.tab, out, partial, \
                     bits)
    if (m.dim == 2) {
      if (kfac) { if (pack) INSP_P1(2, true, true, false); else INSP_P1(2, true, false, false); }
      else { if (pack) INSP_P1(2, false, true, false); else INSP_P1(2, false, false, false); }
    } else if (m.Ka) {
      if (kfac) { if (pack) INSP_P1(3, true, true, true); else INSP_P1(3, true, false, true); }
      else { if (pack) INSP_P1(3, false, true, true); else INSP_P1(3, false, false, true); }
    } else {
      if (kfac) { if (pack) INSP_P1(3, true, true, false); else INSP_P1(3, true, false, false); }
      else { if (pack) INSP_P1(3, false, true, false); else INSP_P1(3, false, false, false); }
    }
#undef INSP_P1
    launch_check("k_apply_p1");
    return;
  }
  const bool pack = m.nbc != nullptr;
  const int* nw = pack ? m.nbc : m.nbr;
#define INSP_APPLY(V, P)                                                                          \
  INSP_DISPATCH_N(m.dim, l.N,                                                                     \
                  hipLaunchKernelGGL((k_apply<D_, N_, T, V, P>), grid_for(m.E), dim3(256), 0, s,  \
                                     m.E, m.stride, u, tr, m.Kt, l.tauS, nw, m.code, m.bctype,    \
                                     kfac, C, l.tab, out, partial))
  if (kfac) { if (pack) INSP_APPLY(true, true); else INSP_APPLY(true, false); }
  else { if (pack) INSP_APPLY(false, true); else INSP_APPLY(false, false); }
#undef INSP_APPLY
}

template <typename T>
void launch_apply_p1_cells(const MeshView& m, const LevelView& l, const T* u, const T* kfac, T* out,
                           hipStream_t s, double* partial) {
  if (m.dim != 3 || l.N != 1 || !m.Ka || !m.nbc_cells || !l.tau_cells)
    throw std::invalid_argument("cell-by-cell order-1 apply: 3D order 1 with the cell-by-cell metric, "
                                "packed words and penalties");
  if (m.E == 0) return;
  const T* C = static_cast<const T*>(l.C);
  const P1NodeBits& bits = p1_node_bits(3);
#define INSP_P1V(V)                                                                               \
  hipLaunchKernelGGL((k_apply_p1<3, T, V, true, true, true>), grid_for(m.E), dim3(256), 0, s, m.E, \
                     m.stride, u, m.Ka, l.tau_cells, m.nbc_cells, m.code, m.bctype, kfac, C, l.tab, \
                     out, partial, bits)
  if (kfac) INSP_P1V(true);
  else INSP_P1V(false);
#undef INSP_P1V
  launch_check("k_apply_p1");
}
template void launch_apply_p1_cells<double>(const MeshView&, const LevelView&, const double*,
                                            const double*, double*, hipStream_t, double*);
template void launch_apply_p1_cells<float>(const MeshView&, const LevelView&, const float*,
                                           const float*, float*, hipStream_t, double*);

int apply_blocks(int64_t E) { return (int)((E + 255) / 256); }

void launch_apply_slip(const MeshView& m, const LevelView& l, const double* u, const double* tr,
                       const double* kfac, double* out, const SlipFaces<double>& slip, hipStream_t s,
                       double* partial) {
  if (m.E == 0) return;
  const double* C = l.C64;
  INSP_DISPATCH_N(m.dim, l.N,
                  hipLaunchKernelGGL((k_apply_slip<D_, N_>), grid_for(m.E), dim3(256), 0, s, m.E,
                                     m.stride, u, tr, m.Kt, l.tauS, m.nbr, m.code, m.bctype, kfac, C,
                                     l.tab, out, partial, slip));
}

template <typename T>
bool launch_apply_cheb(const MeshView& m, const LevelView& l, const T* u, const T* tr,
                       const T* kfac, const T* f, const T* inv, const T* prev, T* x_next, T c1,
                       T c2, hipStream_t s) {
  if (l.frag && mfma_apply_supported(m.dim, l.N)) return false;
  if (x_next == u) throw std::invalid_argument("fused Chebyshev step: x_next must not alias x");
  if (m.E == 0) return true;
  const T* C = static_cast<const T*>(l.C);
  const bool pack = m.nbc != nullptr;
  const int* nw = pack ? m.nbc : m.nbr;
  if (l.N == 1) {
    const P1NodeBits& bits = p1_node_bits(m.dim);
#define INSP_P1C(D, V, P, A)                                                                      \
  hipLaunchKernelGGL((k_apply_p1_cheb<D, T, V, P, A>), grid_for(m.E), dim3(256), 0, s, m.E,       \
                     m.stride, u, A ? m.Ka : m.Kt, l.tauS, nw, m.code, m.bctype, kfac, C, l.tab, f, \
                     inv, prev, x_next, c1, c2, bits)
    if (m.dim == 2) {
      if (kfac) { if (pack) INSP_P1C(2, true, true, false); else INSP_P1C(2, true, false, false); }
      else { if (pack) INSP_P1C(2, false, true, false); else INSP_P1C(2, false, false, false); }
    } else if (m.Ka) {
      if (kfac) { if (pack) INSP_P1C(3, true, true, true); else INSP_P1C(3, true, false, true); }
      else { if (pack) INSP_P1C(3, false, true, true); else INSP_P1C(3, false, false, true); }
    } else {
      if (kfac) { if (pack) INSP_P1C(3, true, true, false); else INSP_P1C(3, true, false, false); }
      else { if (pack) INSP_P1C(3, false, true, false); else INSP_P1C(3, false, false, false); }
    }
#undef INSP_P1C
    launch_check("k_apply_p1_cheb");
    return true;
  }
#define INSP_CHEB(V, P)                                                                           \
  INSP_DISPATCH_N(m.dim, l.N,                                                                     \
                  hipLaunchKernelGGL((k_apply_cheb<D_, N_, T, V, P>), grid_for(m.E), dim3(256), 0, \
                                     s, m.E, m.stride, u, tr, m.Kt, l.tauS, nw, m.code, m.bctype, \
                                     kfac, C, l.tab, f, inv, prev, x_next, c1, c2))
  if (kfac) { if (pack) INSP_CHEB(true, true); else INSP_CHEB(true, false); }
  else { if (pack) INSP_CHEB(false, true); else INSP_CHEB(false, false); }
#undef INSP_CHEB
  launch_check("k_apply_cheb");
  return true;
}
template bool launch_apply_cheb<double>(const MeshView&, const LevelView&, const double*,
                                        const double*, const double*, const double*, const double*,
                                        const double*, double*, double, double, hipStream_t);
template bool launch_apply_cheb<float>(const MeshView&, const LevelView&, const float*,
                                       const float*, const float*, const float*, const float*,
                                       const float*, float*, float, float, hipStream_t);

template void launch_trace<double>(const MeshView&, const LevelView&, const double*, const double*,
                                   double*, hipStream_t);
template void launch_trace<float>(const MeshView&, const LevelView&, const float*, const float*,
                                  float*, hipStream_t);
template void launch_apply<double>(const MeshView&, const LevelView&, const double*, const double*,
                                   const double*, double*, hipStream_t, double*);
template void launch_apply<float>(const MeshView&, const LevelView&, const float*, const float*,
                                  const float*, float*, hipStream_t, double*);

void launch_diag(const MeshView& m, const LevelView& l, const double* kfac, double* diag,
                 hipStream_t s, const double* bnormal, int comp) {
  if (m.E == 0) return;
  if (kfac)
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_diag<D_, N_, true>), grid_for(m.E), dim3(256), 0, s, m.E,
                                       m.stride, m.Kt, l.tauS, m.nbr, m.bctype, kfac, l.C64,
                                       l.tab, diag, bnormal, comp));
  else
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_diag<D_, N_, false>), grid_for(m.E), dim3(256), 0, s, m.E,
                                       m.stride, m.Kt, l.tauS, m.nbr, m.bctype, kfac, l.C64,
                                       l.tab, diag, bnormal, comp));
}

void launch_tau(const MeshView& m, const LevelView& l, const double* kfac, int pen_order,
                double* tauS, hipStream_t s) {
  if (m.E == 0) return;
  if (kfac)
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_tau<D_, N_, true>), grid_for(m.E), dim3(256), 0, s, m.E,
                                       m.stride, m.Kt, m.J, m.nbr, m.code, kfac, l.tab, pen_order, tauS));
  else
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_tau<D_, N_, false>), grid_for(m.E), dim3(256), 0, s, m.E,
                                       m.stride, m.Kt, m.J, m.nbr, m.code, kfac, l.tab, pen_order, tauS));
}

void launch_apply_bc(const MeshView& m, const LevelView& l, const double* bc, const double* kfac,
                     double* rhs, hipStream_t s) {
  if (m.E == 0) return;
  if (kfac)
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_apply_bc<D_, N_, true>), grid_for(m.E), dim3(256), 0, s,
                                       m.E, m.stride, m.Kt, m.J, l.tauS, m.nbr, m.bctype, bc,
                                       kfac, l.C64, l.tab, rhs));
  else
    INSP_DISPATCH_N(m.dim, l.N,
                    hipLaunchKernelGGL((k_apply_bc<D_, N_, false>), grid_for(m.E), dim3(256), 0, s,
                                       m.E, m.stride, m.Kt, m.J, l.tauS, m.nbr, m.bctype, bc,
                                       kfac, l.C64, l.tab, rhs));
}

void launch_mass(const MeshView& m, const LevelView& l, const double* in, double* out,
                 hipStream_t s) {
  if (m.E == 0) return;
  INSP_DISPATCH_N(m.dim, l.N,
                  hipLaunchKernelGGL((k_mass<D_, N_>), grid_for(m.E), dim3(256), 0, s, m.E,
                                     m.stride, m.J, l.C64, in, out));
}

namespace {
constexpr int np_of(int dim, int N) {
  return dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
}

template <typename T, int DIM, int NR, int NC>
void matvec_rc(int64_t E, int64_t stride, const T* A, const T* x, T* y, bool add, hipStream_t s) {
  constexpr int R = np_of(DIM, NR), K = np_of(DIM, NC);
  if (add)
    hipLaunchKernelGGL((k_elem_matvec<T, R, K, true>), grid_for(E), dim3(256), 0, s, E, stride, A,
                       x, y);
  else
    hipLaunchKernelGGL((k_elem_matvec<T, R, K, false>), grid_for(E), dim3(256), 0, s, E, stride, A,
                       x, y);
  launch_check("k_elem_matvec");
}

template <typename T, int DIM, int NR>
void matvec_r(int nc, int64_t E, int64_t stride, const T* A, const T* x, T* y, bool add,
              hipStream_t s) {
  switch (nc) {
    case 1: matvec_rc<T, DIM, NR, 1>(E, stride, A, x, y, add, s); break;
    case 2: matvec_rc<T, DIM, NR, 2>(E, stride, A, x, y, add, s); break;
    case 3: matvec_rc<T, DIM, NR, 3>(E, stride, A, x, y, add, s); break;
    case 4: matvec_rc<T, DIM, NR, 4>(E, stride, A, x, y, add, s); break;
    default: throw std::runtime_error("unsupported order");
  }
}

template <typename T, int DIM>
void matvec_d(int nr, int nc, int64_t E, int64_t stride, const T* A, const T* x, T* y, bool add,
              hipStream_t s) {
  switch (nr) {
    case 1: matvec_r<T, DIM, 1>(nc, E, stride, A, x, y, add, s); break;
    case 2: matvec_r<T, DIM, 2>(nc, E, stride, A, x, y, add, s); break;
    case 3: matvec_r<T, DIM, 3>(nc, E, stride, A, x, y, add, s); break;
    case 4: matvec_r<T, DIM, 4>(nc, E, stride, A, x, y, add, s); break;
    default: throw std::runtime_error("unsupported order");
  }
}
}  // namespace

template <typename T>
void launch_elem_matvec(int dim, int rows_order, int cols_order, int64_t E, int64_t stride,
                        const T* A, const T* x, T* y, bool add, hipStream_t s) {
  if (E == 0) return;
  if (dim == 2)
    matvec_d<T, 2>(rows_order, cols_order, E, stride, A, x, y, add, s);
  else
    matvec_d<T, 3>(rows_order, cols_order, E, stride, A, x, y, add, s);
}

template void launch_elem_matvec<double>(int, int, int, int64_t, int64_t, const double*,
                                         const double*, double*, bool, hipStream_t);
template void launch_elem_matvec<float>(int, int, int, int64_t, int64_t, const float*,
                                        const float*, float*, bool, hipStream_t);

// Scratch memory of the allocation-free apply (ins_poisson_mult_noalloc): the face traces of
// one FP64 apply, then staging copies of its input and output when the caller's dat stride is
// not the solver's. Every GPU the library runs on holds its own copy until the process ends;
// calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

double* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
  if (dev < 0 || dev >= kMaxScratchDevices)
    throw std::runtime_error("device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    // the array's name is a host-side stand-in; the device copy's address comes from the runtime
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace insp
