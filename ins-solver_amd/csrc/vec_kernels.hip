#include "vec_kernels.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <initializer_list>
#include <stdexcept>
#include <string>

namespace insp {

namespace {
inline void launch_check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("kernel launch failed: ") + what + ": " + hipGetErrorString(e));
}
}  // namespace

namespace {

constexpr int kBlock = 256;

// Block-wide sum of K values per thread; every thread gets the totals.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  __shared__ double sh[K][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) t += sh[k][i];
    v[k] = t;
  }
  __syncthreads();
}

template <int K>
__device__ __forceinline__ void write_partials(double (&v)[K], double* partial) {
  block_sum<K>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partial[k * kRedBlocks + blockIdx.x] = v[k];
  }
}

// One block adds the partials of each value: 1024 threads, four independent loads in flight per
// thread (thread t owns entries t + 1024 (4j + u)), combined in a fixed order.
constexpr int kFinBlock = 1024;
__global__ void __launch_bounds__(kFinBlock)
k_finalize(int nb, int nvals, const double* __restrict__ partial, double* __restrict__ scal,
           int slot) {
  for (int k = 0; k < nvals; ++k) {
    const double* src = partial + (size_t)k * kRedBlocks;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += 4 * kFinBlock) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = i + u * kFinBlock;
        if (j < nb) a[u] += src[j];
      }
    }
    const double t = block_sum1<kFinBlock>((a[0] + a[1]) + (a[2] + a[3]));
    if (threadIdx.x == 0) scal[slot + k] = t;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dot(int64_t n, const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
      const T* __restrict__ d, int nd, double* __restrict__ partial) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    v[0] += (double)a[i] * (double)b[i];
    if (nd > 1) v[1] += (double)c[i] * (double)d[i];
  }
  write_partials<2>(v, partial);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dot_flex(int64_t n, const T* __restrict__ r, const T* __restrict__ z, const T* __restrict__ zold,
           double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    v[0] += (double)r[i] * ((double)z[i] - (double)zold[i]);
  write_partials<1>(v, partial);
}

// Both inner products of a flexible-CG step in one pass: r.z and r.(z - zold).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_dot_rz_flex(int64_t n, const T* __restrict__ r, const T* __restrict__ z, const T* __restrict__ zold,
              double* __restrict__ partial) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    v[0] += (double)r[i] * (double)z[i];
    v[1] += (double)r[i] * ((double)z[i] - (double)zold[i]);
  }
  write_partials<2>(v, partial);
}

__global__ void __launch_bounds__(kBlock)
k_residual(int64_t n, const double* __restrict__ b, const double* __restrict__ Ax,
           double* __restrict__ r, double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double t = b[i] - Ax[i];
    r[i] = t;
    v[0] += t * t;
  }
  write_partials<1>(v, partial);
}

__device__ __forceinline__ double safe_div(double a, double b) { return b != 0.0 ? a / b : 0.0; }

// True residual in FP64, stored in FP32 for the inner iterations of the mixed solve.
__global__ void __launch_bounds__(kBlock)
k_residual_f32(int64_t n, const double* __restrict__ b, const double* __restrict__ Ax,
               float* __restrict__ r, double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double t = b[i] - Ax[i];
    r[i] = (float)t;
    v[0] += t * t;
  }
  write_partials<1>(v, partial);
}

// x (FP64) += alpha p, r (FP32) -= alpha Ap, |r|^2 accumulated in FP64.
__global__ void __launch_bounds__(kBlock)
k_cg_update_xr_mixed(int64_t n, double* __restrict__ x, float* __restrict__ r,
                     const float* __restrict__ p, const float* __restrict__ Ap,
                     const double* __restrict__ scal, int num, int den, double* __restrict__ partial) {
  const double alpha = safe_div(scal[num], scal[den]);
  const float a = (float)alpha;
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    x[i] = fma(alpha, (double)p[i], x[i]);
    const float ri = fmaf(-a, Ap[i], r[i]);
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
  }
  write_partials<1>(v, partial);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_xr(int64_t n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
               const T* __restrict__ Ap, const T* __restrict__ invdiag, T* __restrict__ z,
               const double* __restrict__ scal, int num, int den, double* __restrict__ partial) {
  const double alpha = safe_div(scal[num], scal[den]);
  const T a = (T)alpha;
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    x[i] = fma(a, p[i], x[i]);
    const T ri = fma(-a, Ap[i], r[i]);
    r[i] = ri;
    if (invdiag) {
      const T zi = invdiag[i] * ri;
      z[i] = zi;
      v[0] += (double)ri * (double)zi;
    } else {
      v[0] += (double)ri * (double)ri;
    }
  }
  write_partials<1>(v, partial);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_p(int64_t n, T* __restrict__ p, const T* __restrict__ z, const double* __restrict__ scal,
              int num, int den) {
  const T beta = (T)safe_div(scal[num], scal[den]);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = fma(beta, p[i], z[i]);
}

// Jacobi-CG carried in z = D^-1 r (r is never stored): z -= alpha D^-1 Ap, and r.z = sum z^2 / D^-1
// over the cells (padding entries hold D^-1 = 0 and z = 0).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_z(int64_t n, T* __restrict__ z, const T* __restrict__ Ap, const T* __restrict__ invdiag,
              const double* __restrict__ scal, int num, int den, double* __restrict__ partial) {
  const T a = (T)safe_div(scal[num], scal[den]);
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T di = invdiag[i];
    const T zi = fma(-a * di, Ap[i], z[i]);
    z[i] = zi;
    const double zd = (double)zi;
    v[0] += di != (T)0 ? zd * zd / (double)di : 0.0;
  }
  write_partials<1>(v, partial);
}

// alpha = scal[anum]/scal[aden], beta = scal[bnum]/scal[anum]: x += alpha p, then p = z + beta p
// (p is left as it is when last is set).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_xp(int64_t n, T* __restrict__ x, T* __restrict__ p, const T* __restrict__ z,
               const double* __restrict__ scal, int anum, int aden, int bnum, bool last) {
  const T a = (T)safe_div(scal[anum], scal[aden]);
  const T beta = (T)safe_div(scal[bnum], scal[anum]);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T pi = p[i];
    x[i] = fma(a, pi, x[i]);
    if (!last) p[i] = fma(beta, pi, z[i]);
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_p(int64_t n, T* __restrict__ p_next, const T* __restrict__ p, const T* __restrict__ z,
              double* __restrict__ scal, int anum, int aden, int bnum, int asave) {
  const T beta = (T)safe_div(scal[bnum], scal[anum]);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p_next[i] = fma(beta, p[i], z[i]);
  // no block reads scal[asave] in this launch
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[asave] = safe_div(scal[anum], scal[aden]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_x2p(int64_t n, T* __restrict__ x, T* __restrict__ p0, const T* __restrict__ p1,
                const T* __restrict__ z, const double* __restrict__ scal, int asave, int anum,
                int aden, int bnum, bool last) {
  const T a0 = (T)scal[asave];
  const T a1 = (T)safe_div(scal[anum], scal[aden]);
  const T beta = (T)safe_div(scal[bnum], scal[anum]);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T q0 = p0[i], q1 = p1[i];
    x[i] = fma(a1, q1, fma(a0, q0, x[i]));
    if (!last) p0[i] = fma(beta, q1, z[i]);
  }
}

// Two entries per thread with 16-byte (FP64) or 8-byte (FP32) accesses: the direction updates of
// the coarse CG, whose vectors are even-length and aligned. Same arithmetic per entry as the
// one-entry kernels.
template <typename T> struct Pair2;
template <> struct Pair2<double> { using type = double2; };
template <> struct Pair2<float> { using type = float2; };

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_p_pair(int64_t n2, T* __restrict__ p_next, const T* __restrict__ p, const T* __restrict__ z,
                   double* __restrict__ scal, int anum, int aden, int bnum, int asave) {
  using V = typename Pair2<T>::type;
  const T beta = (T)safe_div(scal[bnum], scal[anum]);
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i < n2) {
    const V pi = reinterpret_cast<const V*>(p)[i], zi = reinterpret_cast<const V*>(z)[i];
    V o;
    o.x = fma(beta, pi.x, zi.x);
    o.y = fma(beta, pi.y, zi.y);
    reinterpret_cast<V*>(p_next)[i] = o;
  }
  // no block reads scal[asave] in this launch
  if (blockIdx.x == 0 && threadIdx.x == 0) scal[asave] = safe_div(scal[anum], scal[aden]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_x2p_pair(int64_t n2, T* __restrict__ x, T* __restrict__ p0, const T* __restrict__ p1,
                     const T* __restrict__ z, const double* __restrict__ scal, int asave, int anum,
                     int aden, int bnum, bool last) {
  using V = typename Pair2<T>::type;
  const T a0 = (T)scal[asave];
  const T a1 = (T)safe_div(scal[anum], scal[aden]);
  const T beta = (T)safe_div(scal[bnum], scal[anum]);
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= n2) return;
  const V q0 = reinterpret_cast<const V*>(p0)[i], q1 = reinterpret_cast<const V*>(p1)[i];
  V xi = reinterpret_cast<V*>(x)[i];
  xi.x = fma(a1, q1.x, fma(a0, q0.x, xi.x));
  xi.y = fma(a1, q1.y, fma(a0, q0.y, xi.y));
  reinterpret_cast<V*>(x)[i] = xi;
  if (!last) {
    const V zi = reinterpret_cast<const V*>(z)[i];
    V o;
    o.x = fma(beta, q1.x, zi.x);
    o.y = fma(beta, q1.y, zi.y);
    reinterpret_cast<V*>(p0)[i] = o;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cg_update_z_pair(int64_t n2, T* __restrict__ z, const T* __restrict__ Ap,
                   const T* __restrict__ invdiag, const double* __restrict__ scal, int num, int den,
                   double* __restrict__ partial) {
  using V = typename Pair2<T>::type;
  const T a = (T)safe_div(scal[num], scal[den]);
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n2; i += (int64_t)gridDim.x * kBlock) {
    const V di = reinterpret_cast<const V*>(invdiag)[i], api = reinterpret_cast<const V*>(Ap)[i];
    V zi = reinterpret_cast<V*>(z)[i];
    zi.x = fma(-a * di.x, api.x, zi.x);
    zi.y = fma(-a * di.y, api.y, zi.y);
    reinterpret_cast<V*>(z)[i] = zi;
    const double zx = (double)zi.x, zy = (double)zi.y;
    v[0] += di.x != (T)0 ? zx * zx / (double)di.x : 0.0;
    v[0] += di.y != (T)0 ? zy * zy / (double)di.y : 0.0;
  }
  write_partials<1>(v, partial);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_jacobi(int64_t n, const T* __restrict__ r, const T* __restrict__ invdiag, T* __restrict__ z) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    z[i] = invdiag[i] * r[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cheb_step(int64_t n, const T* __restrict__ f, const T* __restrict__ Ax,
            const T* __restrict__ invdiag, T* __restrict__ d, T* __restrict__ x, T c1, T c2,
            bool x_zero) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T res = invdiag[i] * (Ax ? f[i] - Ax[i] : f[i]);
    const T di = c1 == (T)0 ? c2 * res : fma(c1, d[i], c2 * res);
    if (d) d[i] = di;
    x[i] = x_zero ? di : x[i] + di;
  }
}

// Three-term form: x_next = x + c1 (x - prev) + c2 D^-1 (f - Ax), prev not read when c1 = 0 and
// zero when null; x_next may be prev (each entry is read before it is written).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cheb_step3(int64_t n, const T* __restrict__ f, const T* __restrict__ Ax,
             const T* __restrict__ invdiag, const T* prev, const T* __restrict__ x, T* x_next, T c1,
             T c2) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const T xi = x[i];
    T di = c2 * (invdiag[i] * (f[i] - Ax[i]));
    if (c1 != (T)0) di = fma(c1, prev ? xi - prev[i] : xi, di);
    x_next[i] = xi + di;
  }
}

// Rows (v[i*stride + e]) to cell by cell (v[e*Np + i]) and back, over the whole Np*stride range:
// entries of padding cells, and the cell-by-cell tail past Np*E, are written as zero.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_rows_to_cells(int64_t E, int64_t stride, int Np, const T* __restrict__ src, T* __restrict__ dst) {
  const int64_t o = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (o >= Np * stride) return;
  const int64_t e = o / Np;
  const int i = (int)(o - e * Np);
  dst[o] = e < E ? src[i * stride + e] : (T)0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cells_to_rows(int64_t E, int64_t stride, int Np, const T* __restrict__ src, T* __restrict__ dst) {
  const int64_t o = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (o >= Np * stride) return;
  const int i = (int)(o / stride);
  const int64_t e = o - i * stride;
  dst[o] = e < E ? src[e * Np + i] : (T)0;
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kBlock)
k_convert(int64_t n, const Tin* __restrict__ x, Tout* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = (Tout)x[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill(int64_t n, T* __restrict__ x, T v) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    x[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_axpby(int64_t n, T a, const T* __restrict__ x, T b, T* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = fma(a, x[i], b * y[i]);
}

// 16-byte accesses for the smoother's streams: two FP64 or four FP32 entries per thread, same
// arithmetic per entry as the one-entry kernels.
template <typename T> struct Vec16;
template <> struct Vec16<double> { using type = double2; static constexpr int n = 2; };
template <> struct Vec16<float> { using type = float4; static constexpr int n = 4; };

template <typename T>
__device__ __forceinline__ void ld16(const T* p, int64_t i, T (&o)[Vec16<T>::n]) {
  const typename Vec16<T>::type v = reinterpret_cast<const typename Vec16<T>::type*>(p)[i];
  __builtin_memcpy(o, &v, 16);
}
template <typename T>
__device__ __forceinline__ void st16(T* p, int64_t i, const T (&o)[Vec16<T>::n]) {
  typename Vec16<T>::type v;
  __builtin_memcpy(&v, o, 16);
  reinterpret_cast<typename Vec16<T>::type*>(p)[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_axpby16(int64_t n16, T a, const T* __restrict__ x, T b, T* __restrict__ y) {
  constexpr int W = Vec16<T>::n;
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= n16) return;
  T xv[W], yv[W];
  ld16(x, i, xv);
  ld16(y, i, yv);
#pragma unroll
  for (int k = 0; k < W; ++k) yv[k] = fma(a, xv[k], b * yv[k]);
  st16(y, i, yv);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_cheb_step3_16(int64_t n16, const T* __restrict__ f, const T* __restrict__ Ax,
                const T* __restrict__ invdiag, const T* prev, const T* __restrict__ x, T* x_next,
                T c1, T c2) {
  constexpr int W = Vec16<T>::n;
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= n16) return;
  T fv[W], av[W], dv[W], xv[W], pv[W], o[W];
  ld16(f, i, fv);
  ld16(Ax, i, av);
  ld16(invdiag, i, dv);
  ld16(x, i, xv);
  const bool rd = c1 != (T)0 && prev;
  if (rd) ld16(prev, i, pv);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    T di = c2 * (dv[k] * (fv[k] - av[k]));
    if (c1 != (T)0) di = fma(c1, rd ? xv[k] - pv[k] : xv[k], di);
    o[k] = xv[k] + di;
  }
  st16(x_next, i, o);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_sub_mean(int rows, int64_t E, int64_t stride, T* __restrict__ x, const double* __restrict__ scal,
           int slot, double count) {
  const T mean = (T)(scal[slot] / count);
  const int64_t n = (int64_t)rows * stride;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    if (i % stride < E) x[i] -= mean;
}

__global__ void __launch_bounds__(kBlock)
k_rand_fill(int rows, int64_t E, int64_t stride, uint64_t seed, double* __restrict__ x) {
  const int64_t n = (int64_t)rows * stride;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    // splitmix64 of (seed, row * E + cell): the values do not depend on the row padding
    const int64_t row = i / stride, cell = i - row * stride;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)row * (uint64_t)E + (uint64_t)cell + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * 0x1.0p-53;
    x[i] = cell < E ? 2.0 * u - 1.0 : 0.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_vsum(int64_t n, const T* __restrict__ x, double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    v[0] += (double)x[i];
  write_partials<1>(v, partial);
}

__global__ void __launch_bounds__(kBlock)
k_soa_from_aos(int rows, int64_t E, int64_t stride, const double* __restrict__ aos,
               double* __restrict__ soa) {
  const int64_t n = (int64_t)rows * stride;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t e = i % stride, r = i / stride;
    soa[i] = e < E ? aos[e * rows + r] : 0.0;
  }
}

__global__ void __launch_bounds__(kBlock)
k_aos_from_soa(int rows, int64_t E, int64_t stride, const double* __restrict__ soa,
               double* __restrict__ aos) {
  const int64_t n = (int64_t)rows * E;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t e = i / rows, r = i % rows;
    aos[i] = soa[r * stride + e];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_zero_padding(int rows, int64_t E, int64_t stride, T* __restrict__ x) {
  const int64_t pad = stride - E;
  const int64_t n = (int64_t)rows * pad;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    x[(i / pad) * stride + E + i % pad] = (T)0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_reciprocal(int rows, int64_t E, int64_t stride, const double* __restrict__ diag, T* __restrict__ inv) {
  const int64_t n = (int64_t)rows * stride;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    inv[i] = (i % stride < E) ? (T)(1.0 / diag[i]) : (T)0;
}

__global__ void k_scal_copy(double* scal, int dst, int src) {
  if (threadIdx.x == 0) scal[dst] = scal[src];
}

// Vectors that the pair kernels may take: an even length and every base aligned to two entries.
template <typename T>
bool pairs_ok(int64_t n, std::initializer_list<const void*> ptrs) {
  if (n & 1) return false;
  for (const void* q : ptrs)
    if (reinterpret_cast<uintptr_t>(q) % (2 * sizeof(T)) != 0) return false;
  return true;
}

// Vectors that the 16-byte kernels may take: a length that is a multiple of the entries per
// access and every base (null ones aside) 16-byte aligned.
template <typename T>
bool vec16_ok(int64_t n, std::initializer_list<const void*> ptrs) {
  if (n % (16 / (int64_t)sizeof(T)) != 0) return false;
  for (const void* q : ptrs)
    if (reinterpret_cast<uintptr_t>(q) % 16 != 0) return false;
  return true;
}

// Streams take one element per thread over a grid that covers the vector: on an MI355X that
// streams at 6.1 TB/s where grid-stride loops over a few thousand blocks reach 4.4-4.9.
int grid_stream(int64_t n) {
  const int64_t b = (n + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, INT32_MAX));
}

}  // namespace

// Reductions: about 4 elements per thread, at most kRedBlocks partials.
int red_blocks(int64_t n) {
  const int64_t b = (n + kBlock * 4 - 1) / (kBlock * 4);
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kRedBlocks));
}

void finalize(int nb, int nvals, const double* partial, double* scal, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, nvals, partial, scal, slot);
  launch_check("k_finalize");
}

template <typename T>
void dot(int64_t n, const T* a, const T* b, const T* c, const T* d, int nd, double* scal, int slot,
         RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL((k_dot<T>), dim3(nb), dim3(kBlock), 0, s, n, a, b, c, d, nd, rs.partial); launch_check("k_dot");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, nd, rs.partial, scal, slot); launch_check("k_finalize");
}

template <typename T>
void dot_flex(int64_t n, const T* r, const T* z, const T* zold, double* scal, int slot,
              RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL((k_dot_flex<T>), dim3(nb), dim3(kBlock), 0, s, n, r, z, zold, rs.partial); launch_check("k_dot_flex");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, slot); launch_check("k_finalize");
}

template <typename T>
void dot_rz_flex(int64_t n, const T* r, const T* z, const T* zold, double* scal, int slot,
                 RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL((k_dot_rz_flex<T>), dim3(nb), dim3(kBlock), 0, s, n, r, z, zold, rs.partial); launch_check("k_dot_rz_flex");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 2, rs.partial, scal, slot); launch_check("k_finalize");
}

void residual(int64_t n, const double* b, const double* Ax, double* r, double* scal, int slot,
              RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(kBlock), 0, s, n, b, Ax, r, rs.partial); launch_check("k_residual");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, slot); launch_check("k_finalize");
}

void residual_f32(int64_t n, const double* b, const double* Ax, float* r, double* scal, int slot,
                  RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_residual_f32, dim3(nb), dim3(kBlock), 0, s, n, b, Ax, r, rs.partial); launch_check("k_residual_f32");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, slot); launch_check("k_finalize");
}

void cg_update_xr_mixed(int64_t n, double* x, float* r, const float* p, const float* Ap, double* scal,
                        int num, int den, int out, RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_cg_update_xr_mixed, dim3(nb), dim3(kBlock), 0, s, n, x, r, p, Ap, scal, num, den,
                     rs.partial); launch_check("k_cg_update_xr_mixed");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, out); launch_check("k_finalize");
}

template <typename T>
void cg_update_xr(int64_t n, T* x, T* r, const T* p, const T* Ap, const T* invdiag, T* z,
                  double* scal, int num, int den, int out, RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL((k_cg_update_xr<T>), dim3(nb), dim3(kBlock), 0, s, n, x, r, p, Ap, invdiag, z,
                     scal, num, den, rs.partial); launch_check("k_cg_update_xr");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, out); launch_check("k_finalize");
}

template <typename T>
void cg_update_p(int64_t n, T* p, const T* z, const double* scal, int num, int den, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_update_p<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, p, z, scal,
                     num, den); launch_check("k_cg_update_p");
}

template <typename T>
void cg_update_z(int64_t n, T* z, const T* Ap, const T* invdiag, double* scal, int num, int den,
                 int out, RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  if (pairs_ok<T>(n, {z, Ap, invdiag})) {
    hipLaunchKernelGGL((k_cg_update_z_pair<T>), dim3(nb), dim3(kBlock), 0, s, n / 2, z, Ap, invdiag,
                       scal, num, den, rs.partial); launch_check("k_cg_update_z_pair");
  } else {
    hipLaunchKernelGGL((k_cg_update_z<T>), dim3(nb), dim3(kBlock), 0, s, n, z, Ap, invdiag, scal, num,
                       den, rs.partial); launch_check("k_cg_update_z");
  }
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, out); launch_check("k_finalize");
}

template <typename T>
void cg_update_xp(int64_t n, T* x, T* p, const T* z, const double* scal, int anum, int aden,
                  int bnum, bool last, hipStream_t s) {
  hipLaunchKernelGGL((k_cg_update_xp<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, x, p, z, scal,
                     anum, aden, bnum, last); launch_check("k_cg_update_xp");
}

template <typename T>
void cg_update_p(int64_t n, T* p_next, const T* p, const T* z, double* scal, int anum, int aden,
                 int bnum, int asave, hipStream_t s) {
  if (pairs_ok<T>(n, {p_next, p, z})) {
    hipLaunchKernelGGL((k_cg_update_p_pair<T>), dim3(grid_stream(n / 2)), dim3(kBlock), 0, s, n / 2,
                       p_next, p, z, scal, anum, aden, bnum, asave); launch_check("k_cg_update_p_pair");
    return;
  }
  hipLaunchKernelGGL((k_cg_update_p<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, p_next, p, z, scal,
                     anum, aden, bnum, asave); launch_check("k_cg_update_p");
}

template <typename T>
void cg_update_x2p(int64_t n, T* x, T* p0, const T* p1, const T* z, const double* scal, int asave,
                   int anum, int aden, int bnum, bool last, hipStream_t s) {
  if (pairs_ok<T>(n, {x, p0, p1, z})) {
    hipLaunchKernelGGL((k_cg_update_x2p_pair<T>), dim3(grid_stream(n / 2)), dim3(kBlock), 0, s, n / 2,
                       x, p0, p1, z, scal, asave, anum, aden, bnum, last);
    launch_check("k_cg_update_x2p_pair");
    return;
  }
  hipLaunchKernelGGL((k_cg_update_x2p<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, x, p0, p1, z,
                     scal, asave, anum, aden, bnum, last); launch_check("k_cg_update_x2p");
}

template <typename T>
void jacobi(int64_t n, const T* r, const T* invdiag, T* z, hipStream_t s) {
  hipLaunchKernelGGL((k_jacobi<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, r, invdiag, z); launch_check("k_jacobi");
}

template <typename T>
void cheb_step(int64_t n, const T* f, const T* Ax, const T* invdiag, T* d, T* x, T c1, T c2,
               bool x_zero, hipStream_t s) {
  hipLaunchKernelGGL((k_cheb_step<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, f, Ax, invdiag,
                     d, x, c1, c2, x_zero); launch_check("k_cheb_step");
}

template <typename T>
void rows_to_cells(int64_t E, int64_t stride, int Np, const T* src, T* dst, hipStream_t s) {
  if (src == dst) throw std::invalid_argument("rows_to_cells: the copy is out of place");
  hipLaunchKernelGGL((k_rows_to_cells<T>), dim3(grid_stream(Np * stride)), dim3(kBlock), 0, s, E, stride,
                     Np, src, dst); launch_check("k_rows_to_cells");
}

template <typename T>
void cells_to_rows(int64_t E, int64_t stride, int Np, const T* src, T* dst, hipStream_t s) {
  if (src == dst) throw std::invalid_argument("cells_to_rows: the copy is out of place");
  hipLaunchKernelGGL((k_cells_to_rows<T>), dim3(grid_stream(Np * stride)), dim3(kBlock), 0, s, E, stride,
                     Np, src, dst); launch_check("k_cells_to_rows");
}

template <typename T>
void cheb_step3(int64_t n, const T* f, const T* Ax, const T* invdiag, const T* prev, const T* x,
                T* x_next, T c1, T c2, hipStream_t s) {
  if (x_next == x) throw std::invalid_argument("three-term Chebyshev step: x_next must not alias x");
  if (vec16_ok<T>(n, {f, Ax, invdiag, prev, x, x_next})) {
    const int64_t n16 = n / Vec16<T>::n;
    hipLaunchKernelGGL((k_cheb_step3_16<T>), dim3(grid_stream(n16)), dim3(kBlock), 0, s, n16, f, Ax,
                       invdiag, prev, x, x_next, c1, c2); launch_check("k_cheb_step3_16");
    return;
  }
  hipLaunchKernelGGL((k_cheb_step3<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, f, Ax, invdiag,
                     prev, x, x_next, c1, c2); launch_check("k_cheb_step3");
}

template <typename Tin, typename Tout>
void convert(int64_t n, const Tin* x, Tout* y, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, x, y); launch_check("k_convert");
}

template <typename T>
void fill(int64_t n, T* x, T v, hipStream_t s) {
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, x, v); launch_check("k_fill");
}

template <typename T>
void copy(int64_t n, const T* x, T* y, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<T, T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, x, y); launch_check("k_convert");
}

template <typename T>
void axpby(int64_t n, T a, const T* x, T b, T* y, hipStream_t s) {
  if (vec16_ok<T>(n, {x, y})) {
    const int64_t n16 = n / Vec16<T>::n;
    hipLaunchKernelGGL((k_axpby16<T>), dim3(grid_stream(n16)), dim3(kBlock), 0, s, n16, a, x, b, y);
    launch_check("k_axpby16");
    return;
  }
  hipLaunchKernelGGL((k_axpby<T>), dim3(grid_stream(n)), dim3(kBlock), 0, s, n, a, x, b, y); launch_check("k_axpby");
}

template <typename T>
void sub_mean(int rows, int64_t E, int64_t stride, T* x, const double* scal, int slot,
              double count, hipStream_t s) {
  hipLaunchKernelGGL((k_sub_mean<T>), dim3(grid_stream((int64_t)rows * stride)), dim3(kBlock), 0, s,
                     rows, E, stride, x, scal, slot, count); launch_check("k_sub_mean");
}

void rand_fill(int rows, int64_t E, int64_t stride, uint64_t seed, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_rand_fill, dim3(grid_stream((int64_t)rows * stride)), dim3(kBlock), 0, s,
                     rows, E, stride, seed, x); launch_check("k_rand_fill");
}

template <typename T>
void vsum(int64_t n, const T* x, double* scal, int slot, RedScratch rs, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL((k_vsum<T>), dim3(nb), dim3(kBlock), 0, s, n, x, rs.partial); launch_check("k_vsum");
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, s, nb, 1, rs.partial, scal, slot); launch_check("k_finalize");
}

void soa_from_aos(int rows, int64_t E, int64_t stride, const double* aos, double* soa,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_soa_from_aos, dim3(grid_stream((int64_t)rows * stride)), dim3(kBlock), 0, s,
                     rows, E, stride, aos, soa); launch_check("k_soa_from_aos");
}

void aos_from_soa(int rows, int64_t E, int64_t stride, const double* soa, double* aos,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_aos_from_soa, dim3(grid_stream((int64_t)rows * E)), dim3(kBlock), 0, s,
                     rows, E, stride, soa, aos); launch_check("k_aos_from_soa");
}

template <typename T>
void zero_padding(int rows, int64_t E, int64_t stride, T* x, hipStream_t s) {
  if (stride == E) return;
  hipLaunchKernelGGL((k_zero_padding<T>), dim3(grid_stream((int64_t)rows * (stride - E))),
                     dim3(kBlock), 0, s, rows, E, stride, x); launch_check("k_zero_padding");
}

template <typename T>
void reciprocal(int rows, int64_t E, int64_t stride, const double* diag, T* inv, hipStream_t s) {
  hipLaunchKernelGGL((k_reciprocal<T>), dim3(grid_stream((int64_t)rows * stride)), dim3(kBlock), 0,
                     s, rows, E, stride, diag, inv); launch_check("k_reciprocal");
}

void scal_copy(double* scal, int dst, int src, hipStream_t s) {
  hipLaunchKernelGGL(k_scal_copy, dim3(1), dim3(64), 0, s, scal, dst, src); launch_check("k_scal_copy");
}

#define INSP_INST(T)                                                                              \
  template void dot<T>(int64_t, const T*, const T*, const T*, const T*, int, double*, int,        \
                       RedScratch, hipStream_t);                                                  \
  template void dot_flex<T>(int64_t, const T*, const T*, const T*, double*, int, RedScratch,      \
                            hipStream_t);                                                         \
  template void dot_rz_flex<T>(int64_t, const T*, const T*, const T*, double*, int, RedScratch,   \
                               hipStream_t);                                                      \
  template void cg_update_xr<T>(int64_t, T*, T*, const T*, const T*, const T*, T*, double*, int,  \
                                int, int, RedScratch, hipStream_t);                               \
  template void cg_update_p<T>(int64_t, T*, const T*, const double*, int, int, hipStream_t);      \
  template void cg_update_z<T>(int64_t, T*, const T*, const T*, double*, int, int, int,           \
                               RedScratch, hipStream_t);                                          \
  template void cg_update_xp<T>(int64_t, T*, T*, const T*, const double*, int, int, int, bool,    \
                                hipStream_t);                                                     \
  template void cg_update_p<T>(int64_t, T*, const T*, const T*, double*, int, int, int, int,      \
                               hipStream_t);                                                      \
  template void cg_update_x2p<T>(int64_t, T*, T*, const T*, const T*, const double*, int, int,    \
                                 int, int, bool, hipStream_t);                                    \
  template void jacobi<T>(int64_t, const T*, const T*, T*, hipStream_t);                          \
  template void cheb_step<T>(int64_t, const T*, const T*, const T*, T*, T*, T, T, bool,           \
                             hipStream_t);                                                        \
  template void cheb_step3<T>(int64_t, const T*, const T*, const T*, const T*, const T*, T*, T, T, \
                              hipStream_t);                                                       \
  template void rows_to_cells<T>(int64_t, int64_t, int, const T*, T*, hipStream_t);               \
  template void cells_to_rows<T>(int64_t, int64_t, int, const T*, T*, hipStream_t);               \
  template void fill<T>(int64_t, T*, T, hipStream_t);                                             \
  template void copy<T>(int64_t, const T*, T*, hipStream_t);                                      \
  template void axpby<T>(int64_t, T, const T*, T, T*, hipStream_t);                               \
  template void sub_mean<T>(int, int64_t, int64_t, T*, const double*, int, double, hipStream_t); \
  template void vsum<T>(int64_t, const T*, double*, int, RedScratch, hipStream_t);                \
  template void zero_padding<T>(int, int64_t, int64_t, T*, hipStream_t);                          \
  template void reciprocal<T>(int, int64_t, int64_t, const double*, T*, hipStream_t);
INSP_INST(double)
INSP_INST(float)
template void convert<double, float>(int64_t, const double*, float*, hipStream_t);
template void convert<float, double>(int64_t, const float*, double*, hipStream_t);

}  // namespace insp
