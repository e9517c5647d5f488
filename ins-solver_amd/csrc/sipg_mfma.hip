// Element pass of the SIPG operator on the matrix cores (3D, orders 2-4, FP32).
//
// Same operator as apply_cell (sipg_cell.h; reference formulation
// src/solvers/3d/slip_matrix/factor_viscous_matrix.cpp, fvmf_3d_mult_flux.h,
// vmf_3d_mult_avg_jump.h), evaluated as two chained products per tile of 16 cells:
//
//   G   = D4 U                 D4 rows (4i+s): D_s row i for s<3, unit row i for s=3
//   h_r = k * sum_s Kt_rs G_s   pointwise, in registers
//   out = [Bt_r | EMAT_f | W_fr] [h ; F ; beta_fr sig]
//
// v_mfma_f32_16x16x4f32: lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[4(l/16)+i][l%16], i = 0..3. With the four rows of a node next to each other,
// lane (g, c) of m-tile t of G holds the three gradient components of node 4t+g of
// cell c, so h is formed in registers and feeds the second product as its B operand
// without a shuffle (its k-steps run over quads of nodes). The face fluxes are formed
// directly in B layout: the face nodes of each face are padded to quads, k-step
// f*J4+q covers nodes 4q..4q+3 of face f and lane (g, c) computes node 4q+g of cell c.
// Output rows past the last full 16-row tile (Np is
// 10, 20, 35) go through v_mfma_f32_4x4x1f32, whose sixteen 4x4x1 blocks take the
// same B layout: block 4g + c/4 multiplies four output rows by k-slice g, and the
// four k-slices of a cell are added across lanes at the end.
//
// The operator matrices are stored once per order as ready-made A fragments (one
// value per lane per k-step, build_mfma_fragments). A block copies them into LDS
// once and its waves then sweep 16-cell tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "dg_layout.h"
#include "kernels.h"
#include "sipg_cell.h"

namespace insp {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mac16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mac4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

// Element idx of an array addressed by a 32-bit byte offset from a uniform base, so
// that the load uses a scalar base and one offset VGPR (arrays stay below 4 GiB;
// the launcher checks).
template <typename T>
__device__ __forceinline__ T at(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (uint32_t)sizeof(T));
}

// beta_f[s] = sum_r alpha_f[r] Kt[r][s] for a face index known only at run time
// (alpha_f as in face_alpha: -e_t, -e_s, (1,1,1), -e_r); written as selects so that
// no table of alpha values is built in registers.
template <typename T>
__device__ __forceinline__ T face_beta(int f, T k0, T k1, T k2) {
  return f == 0 ? -k2 : f == 1 ? -k1 : f == 2 ? k0 + k1 + k2 : -k0;
}

// Tile geometry of one order.
template <int N>
struct Shape {
  static constexpr OpLayout L = op_layout(3, N);
  static constexpr int Np = L.Np, Nfp = L.Nfp, Nf = L.Nf, NFN = L.Nf * L.Nfp;
  static constexpr int Q = (Np + 3) / 4;        // node quads: m-tiles and k-steps of G = D4 U
  static constexpr int MTF = Np / 16;           // full 16-row output tiles
  static constexpr int R4 = (Np % 16 + 3) / 4;  // 4-row output groups after them
  static constexpr int J4 = (Nfp + 3) / 4;      // k-steps per face (face nodes padded to quads)
  static constexpr int KF = Nf * J4;            // k-steps over the face-node slots
  static constexpr int KS2 = 3 * Q + 4 * KF;    // k-steps of the second product
  static constexpr int F1 = Q * Q;              // fragments of D4
  static constexpr int F2 = (MTF + R4) * KS2;   // fragments of [Bt | EMAT | W]
  static constexpr int NFRAG = F1 + F2;
  // waves per block: a block's fragments stay in LDS, so the largest order runs
  // 16 waves on one block per CU
  static constexpr int WAVES = N == 4 ? 16 : 8;
};

template <int N, bool VAR>
__global__ void __launch_bounds__(Shape<N>::WAVES * 64)
k_apply_mfma(int64_t E, int64_t stride, const float* __restrict__ u, const float* __restrict__ tr,
             const double* __restrict__ Kt, const double* __restrict__ tauS,
             const int* __restrict__ nbr, const int* __restrict__ code,
             const int* __restrict__ bctype, const float* __restrict__ kfac,
             const float* __restrict__ frag, const int* __restrict__ tab, float* __restrict__ out) {
  using S = Shape<N>;
  using T = float;
  constexpr OpLayout L = S::L;
  constexpr int Np = S::Np, Nfp = S::Nfp, Nf = S::Nf, Q = S::Q, MTF = S::MTF, R4 = S::R4,
                J4 = S::J4, KF = S::KF, KS2 = S::KS2, MO = S::MTF + S::R4;
  __shared__ T fr[S::NFRAG * 64];
  // face masks and neighbour node/slot table (entries < 2^12) as 16-bit words
  __shared__ short tl[L.tab_finv];
  {
    static_assert((S::NFRAG * 64 * sizeof(T)) % 16 == 0, "fragments fill whole 16-byte words");
    constexpr int n4 = S::NFRAG * 64 * (int)sizeof(T) / 16;
    const int4* src = reinterpret_cast<const int4*>(frag);
    int4* dst = reinterpret_cast<int4*>(fr);
#pragma unroll 8
    for (int i = threadIdx.x; i < n4; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < L.tab_finv; i += blockDim.x) tl[i] = (short)tab[i];
  }
  __syncthreads();
  const T* f1 = fr;
  const T* f2 = fr + S::F1 * 64;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const uint32_t st = (uint32_t)stride;

  // Persistent waves: blocks are dispatched round-robin over the 8 XCDs; the blocks
  // of XCD x sweep a contiguous range of 16-cell tiles so that face gathers of
  // neighbouring cells hit that XCD's L2.
  const int64_t ntiles = (E + 15) / 16;
  const int nb_all = gridDim.x, b = blockIdx.x;
  const int nx = nb_all < 8 ? nb_all : 8, x = b % nx;  // one range per XCD (fewer if few blocks)
  const int nbx = nb_all / nx + (x < nb_all % nx ? 1 : 0);
  const int64_t t_end = ntiles * (x + 1) / nx;
#pragma unroll 1
  for (int64_t tile = ntiles * x / nx + (int64_t)(b / nx) * S::WAVES + w; tile < t_end;
       tile += (int64_t)nbx * S::WAVES) {
    const int64_t e = tile * 16 + c;
    const bool live = e < E;
    const uint32_t ec = (uint32_t)e;

    // B operand of G = D4 U: lane (g, c) holds u at nodes 4ks+g of cell c
    T ub[Q];
#pragma unroll
    for (int ks = 0; ks < Q; ++ks) {
      const int node = 4 * ks + g;
      ub[ks] = (live && node < Np) ? at(u, node * st + ec) : (T)0;
    }
    T K[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) K[k] = live ? (T)at(Kt, k * st + ec) : (T)0;
    f32x4 acc[MO];
#pragma unroll
    for (int mo = 0; mo < MO; ++mo) acc[mo] = f32x4{0, 0, 0, 0};
    // one k-step of the second product with B value bv
    auto step2 = [&](int ks, T bv) {
#pragma unroll
      for (int mo = 0; mo < MO; ++mo) {
        const T a = f2[(mo * KS2 + ks) * 64 + lane];
        acc[mo] = mo < MTF ? mac16(a, bv, acc[mo]) : mac4(a, bv, acc[mo]);
      }
    };

    // volume: G = D4 U one node quad at a time, h in registers, then Bt_r h_r
#pragma unroll 1
    for (int t = 0; t < Q; ++t) {
      f32x4 G = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < Q; ++ks) G = mac16(f1[(t * Q + ks) * 64 + lane], ub[ks], G);
      const int node = 4 * t + g;
      T kj = 1;
      if constexpr (VAR) kj = (live && node < Np) ? at(kfac, node * st + ec) : (T)0;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        T h = 0;
#pragma unroll
        for (int s = 0; s < 3; ++s) h = fma(K[kidx(3, r, s)], G[s], h);
        step2(r * Q + t, h * kj);
      }
    }

    // faces: k-step f*J4+q covers face f, lane (g, c) forms the flux of its face node
    // 4q+g (slots past Nfp are zero) and the lift by EMAT_f and W_fr follows. The loads go in
    // rounds ahead of the arithmetic: the words, codes and penalties of all faces first, then
    // per face the cell's own and the neighbour's values and traces at every face node (a lane
    // with no term reads valid entries of its own cell and discards them).
    const uint32_t ea = live ? ec : 0u;
    int nbf[Nf], cf[Nf], modef[Nf];
    T tsf[Nf];
#pragma unroll
    for (int f = 0; f < Nf; ++f) {
      nbf[f] = at(nbr, f * st + ea);
      cf[f] = at(code, f * st + ea);
      tsf[f] = (T)at(tauS, f * st + ea);
    }
#pragma unroll
    for (int f = 0; f < Nf; ++f) {
      // 0 interior, 1 Dirichlet, 2 Neumann or no cell: no terms
      int mode = 2;
      if (live) mode = nbf[f] >= 0 ? 0 : (bctype[-nbf[f] - 1] == 0 ? 1 : 2);
      modef[f] = mode;
    }
#pragma unroll 1
    for (int f = 0; f < Nf; ++f) {
      T beta[3];
#pragma unroll
      for (int s = 0; s < 3; ++s) beta[s] = face_beta(f, K[kidx(3, 0, s)], K[kidx(3, 1, s)], K[kidx(3, 2, s)]);
      const int mode = pick(modef, f), nbw = pick(nbf, f);
      const T ts = pick(tsf, f);
      const uint32_t qn = mode == 0 ? (uint32_t)nbw : ea;
      const int tb = L.tab_nbr + (f * Nf * L.nperm + (mode == 0 ? pick(cf, f) : 0)) * Nfp;
      T umq[J4], tmq[J4], upq[J4], tpq[J4], kq[J4];
#pragma unroll
      for (int q = 0; q < J4; ++q) {
        const int j = 4 * q + g, jj = j < Nfp ? j : 0;
        const int p = tl[L.tab_fmask + f * Nfp + jj];
        const int pk = tl[tb + jj];
        umq[q] = at(u, p * st + ea);
        tmq[q] = at(tr, (f * Nfp + jj) * st + ea);
        upq[q] = at(u, (pk >> 6) * st + qn);
        tpq[q] = at(tr, (pk & 63) * st + qn);
        kq[q] = VAR ? at(kfac, p * st + ea) : (T)1;
      }
#pragma unroll
      for (int q = 0; q < J4; ++q) {
        const int j = 4 * q + g, kk = f * J4 + q;
        T F = 0, sg = 0;
        if (j < Nfp && mode != 2) {
          const T um = umq[q], Tm = tmq[q];
          const T up = mode == 0 ? upq[q] : -um;
          const T Tp = mode == 0 ? tpq[q] : -Tm;
          const T jump = um - up;
          F = (T)0.5 * (ts * jump - (Tm - Tp));
          sg = (T)-0.5 * jump * kq[q];
        }
        step2(3 * Q + kk, F);
#pragma unroll
        for (int r = 0; r < 3; ++r) step2(3 * Q + KF + r * KF + kk, sg * beta[r]);
      }
    }

    // full tiles: lane (g, c) holds nodes 16mo + 4g + i of cell c
#pragma unroll
    for (int mo = 0; mo < MTF; ++mo)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (live) out[(16 * mo + 4 * g + i) * st + ec] = acc[mo][i];
    // 4-row groups: add the four k-slices held by lanes c, c+16, c+32, c+48
#pragma unroll
    for (int rg = 0; rg < R4; ++rg)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        T v = acc[MTF + rg][i];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        const int node = 16 * MTF + 4 * rg + i;
        if (live && g == 0 && node < Np) out[node * st + ec] = v;
      }
  }
}

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}

// Blocks of a kernel resident on the whole device at once (cached per order and
// coefficient variant; one device per process).
int resident_blocks(const void* kernel, int threads, int N, int var) {
  static int cache[5][2] = {};
  int& v = cache[N][var];
  if (v == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        cus <= 0 || per_cu <= 0)
      throw std::runtime_error("matrix-core element pass: occupancy query failed");
    v = cus * per_cu;
  }
  return v;
}

template <int N>
int fragment_count() {
  return Shape<N>::NFRAG * 64;
}

template <int N, bool VAR>
void launch_order(const MeshView& m, const LevelView& l, const float* u, const float* tr,
                  const float* kfac, float* out, hipStream_t s) {
  using SH = Shape<N>;
  // 32-bit element offsets: the largest array indexed is the trace, Nf*Nfp rows
  if ((int64_t)SH::Nf * SH::Nfp * m.stride * (int64_t)sizeof(double) >= (int64_t(1) << 32))
    throw std::runtime_error("matrix-core element pass: mesh too large for 32-bit offsets");
  const int64_t groups = (((m.E + 15) / 16) + SH::WAVES - 1) / SH::WAVES;
  const int res = resident_blocks((const void*)k_apply_mfma<N, VAR>, SH::WAVES * 64, N, VAR ? 1 : 0);
  const dim3 grid((unsigned)std::min<int64_t>(groups, res));
  hipLaunchKernelGGL((k_apply_mfma<N, VAR>), grid, dim3(SH::WAVES * 64), 0, s, m.E, m.stride, u, tr,
                     m.Kt, l.tauS, m.nbr, m.code, m.bctype, kfac,
                     static_cast<const float*>(l.frag), l.tab, out);
}

}  // namespace

bool mfma_apply_supported(int dim, int N) { return dim == 3 && N >= 2 && N <= 4; }

std::vector<double> build_mfma_fragments(int dim, int N, const std::vector<double>& C,
                                         const std::vector<int>& tab) {
  if (!mfma_apply_supported(dim, N)) return {};
  const OpLayout L = op_layout(dim, N);
  const int Np = L.Np, Nfp = L.Nfp;
  const int Q = (Np + 3) / 4, MTF = Np / 16, R4 = (Np % 16 + 3) / 4, J4 = (Nfp + 3) / 4;
  const int KF = L.Nf * J4;
  const int KS2 = 3 * Q + 4 * KF;
  const int nfrag = N == 2 ? fragment_count<2>() : N == 3 ? fragment_count<3>() : fragment_count<4>();
  if (nfrag != (Q * Q + (MTF + R4) * KS2) * 64) throw std::logic_error("fragment layout mismatch");
  std::vector<double> fr((size_t)nfrag, 0.0);
  // D4 (16x16x4 A operand): row 16t+m = (node 4t + m/4, component m%4), column 4ks+k
  for (int t = 0; t < Q; ++t)
    for (int ks = 0; ks < Q; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int m = l % 16, k = l / 16;
        const int i = 4 * t + m / 4, s = m % 4, j = 4 * ks + k;
        double v = 0;
        if (i < Np && j < Np) v = s < 3 ? C[L.off_D + (s * Np + i) * Np + j] : (i == j ? 1.0 : 0.0);
        fr[(size_t)(t * Q + ks) * 64 + l] = v;
      }
  // [Bt | EMAT | W]: the output row of lane l is 16mo + l%16 in a full tile and
  // 16 MTF + 4(mo - MTF) + l%4 in a 4-row group; its k index is k-step ks, slice l/16
  auto entry = [&](int i, int ks, int k) -> double {
    if (i >= Np) return 0.0;
    if (ks < 3 * Q) {
      const int r = ks / Q, t = ks % Q, j = 4 * t + k;
      return j < Np ? C[L.off_Bt + (r * Np + j) * Np + i] : 0.0;
    }
    if (ks < 3 * Q + KF) {
      const int kk = ks - 3 * Q, f = kk / J4, j = 4 * (kk % J4) + k;
      const int li = tab[L.tab_finv + f * Np + i];
      return j < Nfp && li >= 0 ? C[L.off_MfT + (f * Nfp + j) * Nfp + li] : 0.0;
    }
    const int idx = ks - 3 * Q - KF, r = idx / KF, kk = idx % KF;
    const int f = kk / J4, j = 4 * (kk % J4) + k;
    return j < Nfp ? C[L.off_Wt + ((f * 3 + r) * Nfp + j) * Np + i] : 0.0;
  };
  const size_t o2 = (size_t)Q * Q * 64;
  for (int mo = 0; mo < MTF + R4; ++mo)
    for (int ks = 0; ks < KS2; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int i = mo < MTF ? 16 * mo + l % 16 : 16 * MTF + 4 * (mo - MTF) + l % 4;
        fr[o2 + (size_t)(mo * KS2 + ks) * 64 + l] = entry(i, ks, l / 16);
      }
  return fr;
}

template <typename T>
void launch_apply_mfma(const MeshView& m, const LevelView& l, const T* u, const T* tr,
                       const T* kfac, T* out, hipStream_t s, double* partial) {
  if constexpr (!std::is_same<T, float>::value) {
    throw std::runtime_error("matrix-core element pass: FP32 only");
  } else {
    if (!mfma_apply_supported(m.dim, l.N) || !l.frag)
      throw std::runtime_error("matrix-core element pass: unsupported order or no fragments");
    if (partial) throw std::runtime_error("matrix-core element pass: no fused dot");
    switch (l.N * 2 + (kfac ? 1 : 0)) {
      case 4: launch_order<2, false>(m, l, u, tr, kfac, out, s); break;
      case 5: launch_order<2, true>(m, l, u, tr, kfac, out, s); break;
      case 6: launch_order<3, false>(m, l, u, tr, kfac, out, s); break;
      case 7: launch_order<3, true>(m, l, u, tr, kfac, out, s); break;
      case 8: launch_order<4, false>(m, l, u, tr, kfac, out, s); break;
      case 9: launch_order<4, true>(m, l, u, tr, kfac, out, s); break;
      default: throw std::runtime_error("unsupported order");
    }
    check("k_apply_mfma");
  }
}

template void launch_apply_mfma<float>(const MeshView&, const LevelView&, const float*, const float*,
                                       const float*, float*, hipStream_t, double*);
template void launch_apply_mfma<double>(const MeshView&, const LevelView&, const double*,
                                        const double*, const double*, double*, hipStream_t, double*);

}  // namespace insp
