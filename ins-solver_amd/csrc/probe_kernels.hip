// Achievable device peaks, measured on the device the library runs on: a streaming
// copy (HBM bandwidth) and back-to-back matrix-core products in FP32 and FP64. The
// measured rooflines of the solver's kernels are stated against these beside AMD's
// published figures.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.h"

namespace insp {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

typedef int i32x4 __attribute__((ext_vector_type(4)));

// One 16-byte element per thread over a grid that covers the buffer: on an MI355X this
// streams faster (6.1 TB/s) than grid-stride loops with several loads in flight (4.4-4.9).
__global__ void __launch_bounds__(256) k_copy16(int64_t n16, const i32x4* __restrict__ src,
                                                i32x4* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

// The same with non-temporal loads and stores (the data is not kept in the caches).
__global__ void __launch_bounds__(256) k_copy16_nt(int64_t n16, const i32x4* __restrict__ src,
                                                   i32x4* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

// Four independent accumulation chains per wave keep the matrix pipe busy.
__global__ void __launch_bounds__(256) k_mfma_f32(int iters, float seed, float* __restrict__ out) {
  const float a = seed + threadIdx.x * 1e-7f, b = seed - threadIdx.x * 1e-7f;
  f32x4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
  for (int i = 0; i < iters; ++i) {
    c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c3, 0, 0, 0);
  }
  const f32x4 s = c0 + c1 + c2 + c3;
  if (s[0] == 12345.0f) out[blockIdx.x] = s[1];  // keeps the chains live
}

// Eight chains in FP64 (longer dependent latency than FP32); the loop runs iters / 2.
__global__ void __launch_bounds__(256) k_mfma_f64(int iters, double seed, double* __restrict__ out) {
  const double a = seed + threadIdx.x * 1e-9, b = seed - threadIdx.x * 1e-9;
  f64x4 c[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) c[k] = f64x4{0, 0, 0, 0};
  for (int i = 0; i < iters / 2; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[k], 0, 0, 0);
  }
  f64x4 s = c[0];
#pragma unroll
  for (int k = 1; k < 8; ++k) s += c[k];
  if (s[0] == 12345.0) out[blockIdx.x] = s[1];
}

// Best of `reps` timed launches of `launch` (ms), after one untimed launch.
template <typename F>
float best_ms(hipStream_t s, int reps, F launch) {
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  launch();
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    launch();
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

void measure_device_peaks(double* copy_gbs, double* mfma_f32_tflops, double* mfma_f64_tflops) {
  hipStream_t s;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  int dev = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
  // copy: 512 MiB read + 512 MiB written per launch
  const int64_t bytes = int64_t(512) << 20, n16 = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  hip_check(hipMalloc(&src, bytes), "hipMalloc");
  hip_check(hipMalloc(&dst, bytes), "hipMalloc");
  hip_check(hipMemsetAsync(src, 0, bytes, s), "hipMemset");
  const unsigned cgrid = (unsigned)((n16 + 255) / 256);
  const float tc0 = best_ms(s, 5, [&] {
    hipLaunchKernelGGL(k_copy16, dim3(cgrid), dim3(256), 0, s, n16, (const i32x4*)src, (i32x4*)dst);
  });
  hip_check(hipGetLastError(), "k_copy16");
  const float tc1 = best_ms(s, 5, [&] {
    hipLaunchKernelGGL(k_copy16_nt, dim3(cgrid), dim3(256), 0, s, n16, (const i32x4*)src,
                       (i32x4*)dst);
  });
  hip_check(hipGetLastError(), "k_copy16_nt");
  const float tc = tc0 < tc1 ? tc0 : tc1;
  if (copy_gbs) *copy_gbs = 2.0 * (double)bytes / (tc * 1e-3) * 1e-9;
  // matrix cores: 8 waves per CU, 4 chains x iters products of 16x16x4 per wave
  const int grid = cus * 2, iters = 4096;
  const double products = (double)grid * 4 /*waves*/ * 4 /*chains*/ * iters;
  const float tf = best_ms(s, 3, [&] {
    hipLaunchKernelGGL(k_mfma_f32, dim3(grid), dim3(256), 0, s, iters, 1.0f, (float*)dst);
  });
  hip_check(hipGetLastError(), "k_mfma_f32");
  if (mfma_f32_tflops) *mfma_f32_tflops = products * 2.0 * 16 * 16 * 4 / (tf * 1e-3) * 1e-12;
  const float td = best_ms(s, 3, [&] {
    hipLaunchKernelGGL(k_mfma_f64, dim3(grid), dim3(256), 0, s, iters, 1.0, (double*)dst);
  });
  hip_check(hipGetLastError(), "k_mfma_f64");
  if (mfma_f64_tflops) *mfma_f64_tflops = products * 2.0 * 16 * 16 * 4 / (td * 1e-3) * 1e-12;
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  (void)hipFree(src);
  (void)hipFree(dst);
  (void)hipStreamDestroy(s);
}

}  // namespace insp
