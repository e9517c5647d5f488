// Per-cell bodies of the SIPG operator kernels (trace pass, element pass, the
// one-pass order-1 element pass), shared by the single-solve and the batched
// kernels. See sipg_kernels.hip for the formulation and its reference citations.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "dg_layout.h"
#include "kernels.h"
#include "reduce.h"

namespace insp {

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X. Give each
// XCD a contiguous range of cells so that neighbour gathers hit its own L2.
__device__ __forceinline__ int64_t xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, rem = nb & 7, x = b & 7, k = b >> 3;
  return (int64_t)(x * q + (x < rem ? x : rem) + k);
}

// Operator tables (matrices and index tables) are read-only for a whole launch and the
// same for every lane: reading them through the constant address space lets the compiler
// fetch them with scalar loads into SGPRs whether the pointer came from a kernel argument
// or from a problem descriptor.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}

// Coefficient at an index: VAR 0 = constant 1, 1 = from kfac, 2 = from kfac when it is
// not null (decided per problem at run time; one code path for both).
template <int VAR, typename T>
__device__ __forceinline__ T kval(const T* __restrict__ kfac, int64_t idx) {
  if constexpr (VAR == 0) return (T)1;
  else if constexpr (VAR == 1) return kfac[idx];
  else return kfac ? kfac[idx] : (T)1;
}

// Trace pass for one cell: T = sJ n.(k grad u) at every face node. Per node h = k K grad u is
// formed once; a face's trace is sum_r alpha_{f,r} h_r with the compile-time face coefficients
// (one row of h, or the sum of its rows on the slanted face), so no per-face coefficient table
// stays live across the node loop.
template <int DIM, int N, typename T, int VAR>
__device__ __forceinline__ void
trace_cell(int64_t e, int64_t stride, const T* __restrict__ u, const double* __restrict__ Kt,
           const T* __restrict__ kfac, cptr<T> __restrict__ C, T* __restrict__ tr) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf;
  constexpr FmaskTable FM = make_fmask(DIM, N);
  T uu[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) uu[m] = u[m * stride + e];
  T K[L.NK];
#pragma unroll
  for (int k = 0; k < L.NK; ++k) K[k] = (T)Kt[k * stride + e];
#pragma unroll
  for (int p = 0; p < Np; ++p) {
    T g[DIM];
#pragma unroll
    for (int s = 0; s < DIM; ++s) {
      T a = 0;
#pragma unroll
      for (int m = 0; m < Np; ++m) a = fma(C[L.off_D + (s * Np + p) * Np + m], uu[m], a);
      g[s] = a;
    }
    const T kp = kval<VAR>(kfac, p * stride + e);
    T h[DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
      T a = 0;
#pragma unroll
      for (int s = 0; s < DIM; ++s) a = fma(K[kidx(DIM, r, s)], g[s], a);
      h[r] = kp * a;
    }
#pragma unroll
    for (int f = 0; f < Nf; ++f)
#pragma unroll
      for (int j = 0; j < Nfp; ++j)
        if (FM.v[f * Nfp + j] == p) {
          T v = 0;
#pragma unroll
          for (int r = 0; r < DIM; ++r)
            if (face_alpha(DIM, f, r) != 0.0) v += (T)face_alpha(DIM, f, r) * h[r];
          tr[(int64_t)(f * Nfp + j) * stride + e] = v;
        }
  }
}

// Where the element pass leaves (A u) at flat index k: stored as it is, or consumed by a
// functor out(k, (A u)_k, u_k).
template <typename T>
__device__ __forceinline__ void store_out(T* __restrict__ o, int64_t k, T acc, T) {
  o[k] = acc;
}
template <typename T, class O>
__device__ __forceinline__ void store_out(const O& o, int64_t k, T acc, T u) {
  o(k, acc, u);
}

// One Chebyshev step taken in place of the store, in three-term form: res = D^-1 (f - A x),
// x_next = x + c1 (x - x_prev) + c2 res (x_prev not read when c1 = 0, zero when null). x_next goes
// to another buffer than x (the pass still reads x); it may be x_prev's, each entry being read
// before it is written by the same thread.
template <typename T>
struct ChebOut {
  const T* f;
  const T* inv;
  const T* prev;
  T* x_next;
  T c1, c2;
  __device__ __forceinline__ void operator()(int64_t k, T acc, T u) const {
    const T res = inv[k] * (f[k] - acc);
    T di = c2 * res;
    if (c1 != (T)0) di = fma(c1, prev ? u - prev[k] : u, di);
    x_next[k] = u + di;
  }
};

// Face nodes per load round of the element pass (a divisor of Nfp): whole faces where the
// registers allow it (2D, and 3D up to order 3 in FP32 and order 2 in FP64). 3D faces of order 4,
// and of order 3 in FP64, go one node at a time: larger rounds cost a wave per SIMD and measured
// slower (64^3x6 tets, FP64 order 3: 731 us per element pass one node at a time, 774 us two at a
// time, 1021 us whole faces; 48^3x6, FP64 order 4: 1116 us one at a time, 1322 us five).
constexpr int face_chunk(int dim, int N, int bytes) {
  const int nfp = dim == 2 ? N + 1 : (N + 1) * (N + 2) / 2;
  if (dim == 3 && (N == 4 || (N == 3 && bytes == 8))) return 1;
  return nfp;
}

// One cell of the element pass; returns u.(A u) over the cell. PACK: nbr holds the packed
// words (neighbour << 5) | code, -(bface+1) on the boundary, and code is not read.
template <int DIM, int N, typename T, int VAR, bool PACK = false, class Out = T*,
          class Slip = NoSlip>
__device__ __forceinline__ double
apply_cell(int64_t e, int64_t stride, const T* __restrict__ u, const T* __restrict__ tr,
           const double* __restrict__ Kt, const double* __restrict__ tauS,
           const int* __restrict__ nbr, const int* __restrict__ code,
           const int* __restrict__ bctype, const T* __restrict__ kfac, cptr<T> __restrict__ C,
           cptr<int> __restrict__ tab, Out out, const Slip& slip = Slip{}) {
  constexpr OpLayout L = op_layout(DIM, N);
  constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf;
  constexpr FmaskTable FM = make_fmask(DIM, N);
  T uu[Np], acc[Np];
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    uu[m] = u[m * stride + e];
    acc[m] = 0;
  }
  T K[L.NK];
#pragma unroll
  for (int k = 0; k < L.NK; ++k) K[k] = (T)Kt[k * stride + e];
  int nbf[Nf], cf[Nf];
  T tsf[Nf];
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
    if constexpr (PACK) {
      const int w = nbr[f * stride + e];
      nbf[f] = w >= 0 ? w >> 5 : w;
      cf[f] = w & 31;
    } else {
      nbf[f] = nbr[f * stride + e];
      cf[f] = nbf[f] >= 0 ? code[f * stride + e] : 0;
    }
    tsf[f] = (T)tauS[f * stride + e];
  }

  // Volume term, one node column at a time.
  for (int j = 0; j < Np; ++j) {
    T g[DIM];
#pragma unroll
    for (int s = 0; s < DIM; ++s) {
      T a = 0;
#pragma unroll
      for (int m = 0; m < Np; ++m) a = fma(C[L.off_D + (s * Np + j) * Np + m], uu[m], a);
      g[s] = a;
    }
    T kj = 1;
    kj = kval<VAR>(kfac, j * stride + e);
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
      T h = 0;
#pragma unroll
      for (int s = 0; s < DIM; ++s) h = fma(K[kidx(DIM, r, s)], g[s], h);
      h *= kj;
#pragma unroll
      for (int i = 0; i < Np; ++i) acc[i] = fma(C[L.off_Bt + (r * Np + j) * Np + i], h, acc[i]);
    }
  }

  // Face terms. The face nodes are taken in chunks of CH; the loads of a chunk are issued in
  // two rounds before its arithmetic: the cell's own traces with the node map of the shared
  // face, then the neighbour's values and traces at the mapped nodes. A boundary face reads
  // the cell's own entries in the second round and does not use them. With one chunk per face
  // the loops unroll completely and the cell's own face-node values come from registers.
  constexpr int CH = face_chunk(DIM, N, (int)sizeof(T));
  static_assert(Nfp % CH == 0, "face chunks");
  constexpr bool whole = CH == Nfp;
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
    T beta[DIM];
#pragma unroll
    for (int s = 0; s < DIM; ++s) {
      T b = 0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) b += (T)face_alpha(DIM, f, r) * K[kidx(DIM, r, s)];
      beta[s] = b;
    }
    const int nb = nbf[f];
    int mode = 0;  // 0 interior, 1 Dirichlet, 2 Neumann (or other: no terms), 3 slip
    if (nb < 0) {
      const int bt = bctype[-nb - 1];
      mode = bt == 0 ? 1 : (Slip::enabled && bt == 2 ? 3 : 2);
    }
    if (mode == 2) continue;
    const int64_t q = nb >= 0 ? (int64_t)nb : e;
    const int tbase = L.tab_nbr + (f * Nf * L.nperm + (nb >= 0 ? cf[f] : 0)) * Nfp;
    const T ts = tsf[f];
    // FP32 order 3 in 3D lifts the symmetry term through the face mass and the rows of D_r
    // (fewer coefficient loads, 0.47 vs 0.51 ms on 262k cells); elsewhere the extra live
    // registers cost more than the loads saved (FP64 order 4 spills), so the D_r^T EMAT_f
    // table is read per face node.
    constexpr bool fold = sizeof(T) == 4 && DIM == 3 && N == 3;
    T sv[Nfp];
#pragma unroll
    for (int i = 0; i < Nfp; ++i) sv[i] = 0;
    auto chunk = [&](int j0) {
      T Tm[CH], up[CH], Tp[CH], um[CH];
      int pk[CH];
#pragma unroll
      for (int jj = 0; jj < CH; ++jj) {
        const int j = j0 + jj;
        Tm[jj] = tr[(int64_t)(f * Nfp + j) * stride + e];
        pk[jj] = tab[tbase + j];
        if constexpr (whole) um[jj] = uu[FM.v[f * Nfp + j]];
        else um[jj] = u[tab[L.tab_fmask + f * Nfp + j] * stride + e];
      }
#pragma unroll
      for (int jj = 0; jj < CH; ++jj) {
        up[jj] = u[(int64_t)(pk[jj] >> 6) * stride + q];
        Tp[jj] = tr[(int64_t)(pk[jj] & 63) * stride + q];
      }
#pragma unroll
      for (int jj = 0; jj < CH; ++jj) {
        const int j = j0 + jj;
        const int p = whole ? FM.v[f * Nfp + j] : tab[L.tab_fmask + f * Nfp + j];
        T umj = um[jj];
        T tm = Tm[jj];
        if constexpr (Slip::enabled) {
          if (mode == 3) {
            const double* n = slip.bn + (int64_t)(-nb - 1) * DIM;
            T su = 0, st = 0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              su = fma((T)n[d], slip.u0[d * slip.cstride + p * stride + e], su);
              st = fma((T)n[d], slip.tr0[d * slip.tstride + (int64_t)(f * Nfp + j) * stride + e], st);
            }
            umj = (T)n[slip.comp] * su;
            tm = (T)n[slip.comp] * st;
          }
        }
        const T upj = mode == 0 ? up[jj] : -umj;
        const T tpj = mode == 0 ? Tp[jj] : -tm;
        const T jump = umj - upj;
        const T F = (T)0.5 * (ts * jump - (tm - tpj));
        T sig = (T)-0.5 * jump;
        sig *= kval<VAR>(kfac, p * stride + e);
#pragma unroll
        for (int i = 0; i < Nfp; ++i) {
          const T mf = C[L.off_MfT + (f * Nfp + j) * Nfp + i];
          acc[FM.v[f * Nfp + i]] = fma(mf, F, acc[FM.v[f * Nfp + i]]);
          if constexpr (fold) sv[i] = fma(mf, sig, sv[i]);
        }
        if constexpr (!fold) {
#pragma unroll
          for (int r = 0; r < DIM; ++r) {
            const T bs = beta[r] * sig;
#pragma unroll
            for (int i = 0; i < Np; ++i)
              acc[i] = fma(C[L.off_Wt + ((f * DIM + r) * Nfp + j) * Np + i], bs, acc[i]);
          }
        }
      }
    };
    if constexpr (whole) {
      chunk(0);
    } else {
#pragma unroll 1
      for (int j0 = 0; j0 < Nfp; j0 += CH) chunk(j0);
    }
    // symmetry term sum_r beta_r D_r^T (EMAT_f sig): the face mass first, then the rows
    // of D_r at the face nodes
    if constexpr (fold) {
#pragma unroll
      for (int l = 0; l < Nfp; ++l) {
        const int m = FM.v[f * Nfp + l];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
          const T w = beta[r] * sv[l];
#pragma unroll
          for (int i = 0; i < Np; ++i) acc[i] = fma(C[L.off_D + (r * Np + m) * Np + i], w, acc[i]);
        }
      }
    }
  }
  double dotv = 0.0;
#pragma unroll
  for (int i = 0; i < Np; ++i) {
    store_out(out, (int64_t)i * stride + e, acc[i], uu[i]);
    dotv += (double)uu[i] * (double)acc[i];
  }
  return dotv;
}

// Order 1 in one pass: the gradient of a P1 cell is constant, so the neighbour's
// flux trace k n.grad(u) is recomputed from its Np nodal values and metric instead
// of being stored by a trace pass. Same operator as apply_cell. The face coefficients
// alpha_{f,r} (face_alpha) are 1 for every r on the slanted face f = DIM-1 and -1 for one r
// on the others: r = DIM-1-f before the slanted face, DIM-f after it.
constexpr bool face_alpha_is_rows(int dim) {
  for (int f = 0; f <= dim; ++f)
    for (int r = 0; r < dim; ++r) {
      const int row = f < dim - 1 ? dim - 1 - f : dim - f;
      const double want = f == dim - 1 ? 1.0 : (r == row ? -1.0 : 0.0);
      if (face_alpha(dim, f, r) != want) return false;
    }
  return true;
}
static_assert(face_alpha_is_rows(2) && face_alpha_is_rows(3), "face coefficient rows");

// Entry i (0 <= i < n) of a small register array, selected by a chain of compares. The chain
// starts from zero rather than from v[0], so that no select has two array entries as its
// operands: the compiler would turn such a select into a load through a selected address and
// move the array to scratch memory.
template <int n, typename T>
__device__ __forceinline__ T pick(const T (&v)[n], int i) {
  T r = (T)0;
#pragma unroll
  for (int k = 0; k < n; ++k) r = i == k ? v[k] : r;
  return r;
}

// PACK: nbr holds one word per face, (neighbour << 5) | code inside the mesh and the
// boundary index -(bface+1) on the boundary, and code is not read.
//
// Loads are issued in two rounds, so that a wave waits for memory twice per cell instead of
// several times per face: first everything addressed by the cell index (neighbour words,
// nodal values, metric, penalties), then everything addressed by a neighbour index (the
// neighbour's nodal values and whole metric, and the node map of the shared face). A
// boundary face reads the cell's own entries in the second round and does not use them.
// Where the order-1 element pass finds the neighbour's node at a face node: in the operator's
// integer table (TabNodes: one load per face node), or in bit fields passed with the launch
// (BitNodes: no load, and no registers held while the other loads are in flight).
struct TabNodes {};
struct BitNodes {
  P1NodeBits b;
};

template <int DIM, typename T, int VAR, bool PACK = false, class Out = T*, class Slip = NoSlip,
          class Nodes = TabNodes, bool KAOS = false, bool VCM = false>
__device__ __forceinline__ double
apply_p1_cell(int64_t e, int64_t stride, const T* __restrict__ u, const double* __restrict__ Kt,
              const double* __restrict__ tauS, const int* __restrict__ nbr,
              const int* __restrict__ code, const int* __restrict__ bctype,
              const T* __restrict__ kfac, cptr<T> __restrict__ C, cptr<int> __restrict__ tab,
              Out out, const Slip& slip = Slip{}, const Nodes& nodes = Nodes{}) {
  constexpr OpLayout L = op_layout(DIM, 1);
  constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf, NK = L.NK;
  constexpr FmaskTable FM = make_fmask(DIM, 1);
  // round 1; VCM: nbr holds the packed words and tauS the penalties cell by cell ([E][4])
  static_assert(!VCM || PACK, "cell-by-cell values: packed words");
  int nb[Nf], c[Nf];
  int wcell[Nf];
  if constexpr (VCM) {
    const int4 w4 = reinterpret_cast<const int4*>(nbr)[e];
    wcell[0] = w4.x; wcell[1] = w4.y; wcell[2] = w4.z; wcell[3] = w4.w;
  }
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
    if constexpr (PACK) {
      const int w = VCM ? wcell[f] : nbr[f * stride + e];
      nb[f] = w >= 0 ? w >> 5 : w;
      c[f] = w >= 0 ? (w & 31) : 0;
    } else {
      nb[f] = nbr[f * stride + e];
      c[f] = nb[f] >= 0 ? code[f * stride + e] : 0;
    }
  }
  // VCM: u and out hold the values cell by cell ([E][Np], 3D: 16 or 32 bytes per cell) instead of
  // in rows; the coefficient stays in rows
  static_assert(!VCM || Np == 4, "cell-by-cell values: 3D order 1");
  auto values = [&](int64_t q, T* dst) {
    if constexpr (VCM) {
      if constexpr (sizeof(T) == 8) {
        const double2* uq = reinterpret_cast<const double2*>(u) + q * 2;
        const double2 a = uq[0], b = uq[1];
        dst[0] = a.x; dst[1] = a.y; dst[2] = b.x; dst[3] = b.y;
      } else {
        const float4 a = reinterpret_cast<const float4*>(u)[q];
        dst[0] = a.x; dst[1] = a.y; dst[2] = a.z; dst[3] = a.w;
      }
    } else {
#pragma unroll
      for (int m = 0; m < Np; ++m) dst[m] = u[m * stride + q];
    }
  };
  T uu[Np], acc[Np], kk[Np];
  values(e, uu);
#pragma unroll
  for (int m = 0; m < Np; ++m) {
    acc[m] = 0;
    kk[m] = kval<VAR>(kfac, m * stride + e);
  }
  // KAOS: Kt holds the metric cell by cell ([E][NK]) instead of in rows
  static_assert(!KAOS || NK % 2 == 0, "the cell-by-cell metric is read in pairs");
  auto metric = [&](int64_t q, T* dst) {
    if constexpr (KAOS) {
      const double2* kq = reinterpret_cast<const double2*>(Kt) + q * (NK / 2);
#pragma unroll
      for (int k = 0; k < NK / 2; ++k) {
        const double2 v = kq[k];
        dst[2 * k] = (T)v.x;
        dst[2 * k + 1] = (T)v.y;
      }
    } else {
#pragma unroll
      for (int k = 0; k < NK; ++k) dst[k] = (T)Kt[k * stride + q];
    }
  };
  T K[NK];
  metric(e, K);
  T ts[Nf];
  if constexpr (VCM) {
    const double2* tq = reinterpret_cast<const double2*>(tauS) + e * 2;
    const double2 t01 = tq[0], t23 = tq[1];
    ts[0] = (T)t01.x; ts[1] = (T)t01.y; ts[2] = (T)t23.x; ts[3] = (T)t23.y;
  } else {
#pragma unroll
    for (int f = 0; f < Nf; ++f) ts[f] = (T)tauS[f * stride + e];
  }
  // round 2
  T un[Nf][Np], kn[Nf][NK], fn_k[Nf][Np];
  int nd[Nf][Nfp], bt[Nf];
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
    const bool in = nb[f] >= 0;
    const int64_t q = in ? (int64_t)nb[f] : e;
    const int cf = in ? c[f] : 0;
    values(q, un[f]);
#pragma unroll
    for (int m = 0; m < Np; ++m) fn_k[f][m] = kval<VAR>(kfac, m * stride + q);
    metric(q, kn[f]);
#pragma unroll
    for (int j = 0; j < Nfp; ++j) {
      if constexpr (std::is_same<Nodes, BitNodes>::value)  // 3D: the same row for every face
        nd[f][j] = (int)((nodes.b.w[DIM == 3 ? 0 : f][j] >> (2 * cf)) & 3ull);
      else
        nd[f][j] = tab[L.tab_nbr + (f * Nf * L.nperm + cf) * Nfp + j] >> 6;
    }
    bt[f] = in ? 0 : bctype[-nb[f] - 1];
  }

  // P1: grad phi_i is constant, so Bt[r][j][i] = dphi_ri m_j and Wt[f][r][j][i] = dphi_ri w_j
  // (m, w: column sums of the mass and face mass). The volume term and every face's symmetry
  // term collapse into one vector V: acc_i += sum_r dphi_ri V_r at the end, with
  // V = (sum_j m_j k_j) h + sum_f beta_f sum_j w_j sig_fj.
  (void)C;
  // c_f: the face mass factor; c_s = -w/2 with the face mass column sum w = c_f (Nfp + 1)
  constexpr T c_m = (T)p1_mass_colsum(DIM), c_f = (T)p1_face_mass_c(DIM),
              c_s = (T)(-0.5 * p1_face_mass_c(DIM) * (Nfp + 1));
  double dotv = 0.0;
  T g[DIM];
#pragma unroll
  for (int s = 0; s < DIM; ++s) g[s] = (T)0.5 * (uu[s + 1] - uu[0]);
  T h[DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    T a = 0;
#pragma unroll
    for (int s = 0; s < DIM; ++s) a = fma(K[kidx(DIM, r, s)], g[s], a);
    h[r] = a;
  }
  T V[DIM];
  {
    T ks = kk[0];
#pragma unroll
    for (int j = 1; j < Np; ++j) ks += kk[j];
    const T mk = c_m * ks;
#pragma unroll
    for (int r = 0; r < DIM; ++r) V[r] = mk * h[r];
  }
#pragma unroll
  for (int f = 0; f < Nf; ++f) {
    T beta[DIM];
    T tm0 = 0;
#pragma unroll
    for (int s = 0; s < DIM; ++s) {
      T b = 0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) b += (T)face_alpha(DIM, f, r) * K[kidx(DIM, r, s)];
      beta[s] = b;
      tm0 = fma(b, g[s], tm0);
    }
    // 0 interior, 1 Dirichlet, 2 Neumann (or other: no terms), 3 slip
    const int mode = nb[f] >= 0 ? 0 : (bt[f] == 0 ? 1 : (Slip::enabled && bt[f] == 2 ? 3 : 2));
    if (mode == 2) continue;
    // the neighbour's flux trace sum_s (sum_r alpha_{fn,r} K_n[r][s]) (grad u_n)_s, constant on
    // a P1 cell; alpha_{fn,r} is selected per lane from the compile-time face table
    T tn0 = 0;
    if (mode == 0) {
      const int fn = c[f] / L.nperm;
      T gn[DIM];
#pragma unroll
      for (int s = 0; s < DIM; ++s) gn[s] = (T)0.5 * (un[f][s + 1] - un[f][0]);
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        T al = (T)face_alpha(DIM, 0, r);
#pragma unroll
        for (int k = 1; k < Nf; ++k) al = fn == k ? (T)face_alpha(DIM, k, r) : al;
        T t = 0;
#pragma unroll
        for (int s = 0; s < DIM; ++s) t = fma(kn[f][kidx(DIM, r, s)], gn[s], t);
        tn0 = fma(al, t, tn0);
      }
    }
    // slip: n_c (n.U) at the face nodes and n_c sum_d n_d (beta.grad u_d), the gradients of
    // the components recomputed from their nodal values (constant on a P1 cell)
    T nc = 0, sm0 = 0;
    const double* bnv = nullptr;
    if constexpr (Slip::enabled) {
      if (mode == 3) {
        bnv = slip.bn + (int64_t)(-nb[f] - 1) * DIM;
        nc = (T)bnv[slip.comp];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const T* ud = slip.u0 + d * slip.cstride;
          const T u0 = ud[e];
          T t = 0;
#pragma unroll
          for (int s2 = 0; s2 < DIM; ++s2) t = fma(beta[s2], (T)0.5 * (ud[(s2 + 1) * stride + e] - u0), t);
          sm0 = fma((T)bnv[d], t, sm0);
        }
      }
    }
    T F[Nfp], Fs = 0, S = 0;
#pragma unroll
    for (int j = 0; j < Nfp; ++j) {
      const int p = FM.v[f * Nfp + j];
      T um = uu[p];
      T Tm = kk[p] * tm0;
      if constexpr (Slip::enabled) {
        if (mode == 3) {
          T su = 0;
#pragma unroll
          for (int d = 0; d < DIM; ++d) su = fma((T)bnv[d], slip.u0[d * slip.cstride + p * stride + e], su);
          um = nc * su;
          Tm = kk[p] * nc * sm0;
        }
      }
      T up, Tp;
      if (mode == 0) {
        up = pick(un[f], nd[f][j]);
        Tp = (VAR == 0 ? (T)1 : pick(fn_k[f], nd[f][j])) * tn0;
      } else {
        up = -um;
        Tp = -Tm;
      }
      const T jump = um - up;
      F[j] = (T)0.5 * (ts[f] * jump - (Tm - Tp));
      Fs += F[j];
      S = fma(jump, kk[p], S);
    }
    // face mass c (1 + delta_ij): sum_j M_ij F_j = c (F_i + sum_j F_j); sig_j = -jump_j k_j / 2
#pragma unroll
    for (int i = 0; i < Nfp; ++i) acc[FM.v[f * Nfp + i]] = fma(c_f, F[i] + Fs, acc[FM.v[f * Nfp + i]]);
    const T sw = c_s * S;
#pragma unroll
    for (int r = 0; r < DIM; ++r) V[r] = fma(beta[r], sw, V[r]);
  }
  {
    T v0 = V[0];
#pragma unroll
    for (int r = 1; r < DIM; ++r) v0 += V[r];
    acc[0] = fma((T)-0.5, v0, acc[0]);
#pragma unroll
    for (int r = 0; r < DIM; ++r) acc[r + 1] = fma((T)0.5, V[r], acc[r + 1]);
  }
  if constexpr (VCM) {
    static_assert(std::is_same<Out, T*>::value, "cell-by-cell values: plain output");
    if constexpr (sizeof(T) == 8) {
      double2* oq = reinterpret_cast<double2*>(out) + e * 2;
      oq[0] = make_double2(acc[0], acc[1]);
      oq[1] = make_double2(acc[2], acc[3]);
    } else {
      reinterpret_cast<float4*>(out)[e] = make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < Np; ++i) store_out(out, (int64_t)i * stride + e, acc[i], uu[i]);
  }
#pragma unroll
  for (int i = 0; i < Np; ++i) dotv += (double)uu[i] * (double)acc[i];
  return dotv;
}


}  // namespace insp
