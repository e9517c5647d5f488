// Device kernels of the vertex coarse space: CSR products (a group of lanes per row),
// the dense coarsest solve, vertex transfer to/from the order-1 DG space and
// the probing kernels that assemble the vertex operator.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "amg.h"
#include "reduce.h"

namespace insp {
namespace amg {

namespace {

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("coarse-space kernel launch failed: ") + what + ": " +
                             hipGetErrorString(e));
}

constexpr int kB = 256;
inline unsigned blocks(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

// LANES adjacent lanes share a row: lane l takes the row's nonzeros l, l + LANES, ... (the group's
// loads of val and col are contiguous), then the group adds its partial sums by shuffles in a
// fixed order.
template <typename T, int LANES>
__global__ void __launch_bounds__(kB)
k_spmv(int n, const int* __restrict__ ptr, const int* __restrict__ col, const T* __restrict__ val,
       const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ b, int mode) {
  const int64_t g = blockIdx.x * (int64_t)kB + threadIdx.x;
  const int i = (int)(g / LANES), l = (int)(g % LANES);
  T s = 0;
  if (i < n)
    for (int k = ptr[i] + l; k < ptr[i + 1]; k += LANES) s = fma(val[k], x[col[k]], s);
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LANES);
  if (i >= n || l != 0) return;
  if (mode == 0) y[i] = s;
  else if (mode == 1) y[i] += s;
  else y[i] = b[i] - s;
}

// One block per row of the dense inverse.
template <typename T>
__global__ void __launch_bounds__(kB)
k_dense_matvec(int n, const T* __restrict__ M, const T* __restrict__ x, T* __restrict__ y) {
  const int i = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += kB) s += (double)M[(int64_t)i * n + j] * (double)x[j];
  s = block_sum1<kB>(s);
  if (threadIdx.x == 0) y[i] = (T)s;
}

template <typename T>
__global__ void __launch_bounds__(kB)
k_vertex_restrict(int nv, const int* __restrict__ vptr, const int* __restrict__ vlist,
                  const T* __restrict__ r, T* __restrict__ bv) {
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= nv) return;
  T s = 0;
  for (int k = vptr[v]; k < vptr[v + 1]; ++k) s += r[vlist[k]];
  bv[v] = s;
}

template <typename T>
__global__ void __launch_bounds__(kB)
k_vertex_prolong(int rows, int64_t E, int64_t stride, const int* __restrict__ cv,
                 const T* __restrict__ xv, T* __restrict__ x, int add) {
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= E) return;
  for (int i = 0; i < rows; ++i) {
    const int64_t q = i * stride + e;
    const T v = xv[cv[q]];
    x[q] = add ? x[q] + v : v;
  }
}

__global__ void __launch_bounds__(kB)
k_probe_fill(int nv, const int* __restrict__ colour, int c, double* __restrict__ x) {
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v < nv) x[v] = colour[v] == c ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(kB)
k_probe_extract(int nv, const int* __restrict__ ptr, const int* __restrict__ col,
                const int* __restrict__ colour, int c, const double* __restrict__ y,
                double* __restrict__ val) {
  const int r = blockIdx.x * kB + threadIdx.x;
  if (r >= nv) return;
  for (int k = ptr[r]; k < ptr[r + 1]; ++k)
    if (colour[col[k]] == c) val[k] = y[r];
}

}  // namespace

template <typename T>
void spmv(const DevCsr<T>& A, const T* x, T* y, const T* b, int mode, hipStream_t s) {
  if (A.n == 0) return;
  const unsigned nb = blocks((int64_t)A.n * A.lanes);
#define INSP_SPMV(LN)                                                                             \
  hipLaunchKernelGGL((k_spmv<T, LN>), dim3(nb), dim3(kB), 0, s, A.n, A.ptr, A.col, A.val, x, y, b, mode)
  switch (A.lanes) {
    case 1: INSP_SPMV(1); break;
    case 2: INSP_SPMV(2); break;
    case 4: INSP_SPMV(4); break;
    case 8: INSP_SPMV(8); break;
    case 16: INSP_SPMV(16); break;
    case 32: INSP_SPMV(32); break;
    default: throw std::invalid_argument("CSR product: lanes per row must be 1, 2, 4, 8, 16 or 32");
  }
#undef INSP_SPMV
  check("k_spmv");
}

template <typename T>
void dense_matvec(int n, const T* M, const T* x, T* y, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_dense_matvec<T>), dim3((unsigned)n), dim3(kB), 0, s, n, M, x, y);
  check("k_dense_matvec");
}

template <typename T>
void vertex_restrict(int nv, const int* vptr, const int* vlist, const T* r, T* bv, hipStream_t s) {
  if (nv == 0) return;
  hipLaunchKernelGGL((k_vertex_restrict<T>), dim3(blocks(nv)), dim3(kB), 0, s, nv, vptr, vlist, r, bv);
  check("k_vertex_restrict");
}

template <typename T>
void vertex_prolong(int rows, int64_t E, int64_t stride, const int* cv, const T* xv, T* x, bool add,
                    hipStream_t s) {
  if (E == 0) return;
  hipLaunchKernelGGL((k_vertex_prolong<T>), dim3(blocks(E)), dim3(kB), 0, s, rows, E, stride, cv, xv,
                     x, (int)add);
  check("k_vertex_prolong");
}

void probe_fill(int nv, const int* colour, int c, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_probe_fill, dim3(blocks(nv)), dim3(kB), 0, s, nv, colour, c, x);
  check("k_probe_fill");
}

void probe_extract(int nv, const int* ptr, const int* col, const int* colour, int c, const double* y,
                   double* val, hipStream_t s) {
  hipLaunchKernelGGL(k_probe_extract, dim3(blocks(nv)), dim3(kB), 0, s, nv, ptr, col, colour, c, y, val);
  check("k_probe_extract");
}

#define INSP_AMG_INST(T)                                                                            \
  template void spmv<T>(const DevCsr<T>&, const T*, T*, const T*, int, hipStream_t);                \
  template void dense_matvec<T>(int, const T*, const T*, T*, hipStream_t);                         \
  template void vertex_restrict<T>(int, const int*, const int*, const T*, T*, hipStream_t);        \
  template void vertex_prolong<T>(int, int64_t, int64_t, const int*, const T*, T*, bool, hipStream_t);
INSP_AMG_INST(double)
INSP_AMG_INST(float)

}  // namespace amg
}  // namespace insp
