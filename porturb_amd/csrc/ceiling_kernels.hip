// ceiling_kernels.hip -- the device's own ceilings for the rooflines of this library's kernels (no
// counterpart in the reference): the vector FMA rate in FP64 and in non-packed FP32 (the forms
// every kernel here is compiled to: packed FP32 instructions are switched off), and the HBM rate of
// the plain streaming copy (k_copy) that also calibrates the traffic counters.
#include "pu_common.h"

namespace pu {

// Every thread runs FMA_CHAINS independent multiply-add chains, enough to cover the FMA latency at
// one wave per SIMD; one store per thread keeps the chains live.
constexpr int FMA_CHAINS = 16;

template <class T>
__global__ __launch_bounds__(256) void k_fma_rate(T *out, T a, T b, int iters) {
  T x[FMA_CHAINS];
#pragma unroll
  for (int c = 0; c < FMA_CHAINS; c++) x[c] = (T)(threadIdx.x + c) * (T)1e-3;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < FMA_CHAINS; c++) x[c] = fma(x[c], a, b);
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < FMA_CHAINS; c++) s += x[c];
  out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class T> static int fma_rate(double *tflops) {
  const int blocks = 8192, iters = 4096, reps = 5;
  T *d = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  hipError_t e = hipMalloc(&d, sizeof(T) * (size_t)blocks * 256);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fma_rate<T>, dim3(blocks), dim3(256), 0, nullptr, d, (T)0.999, (T)1e-4, iters);   // warm-up
    e = hipEventRecord(e0, nullptr);
  }
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    hipLaunchKernelGGL(k_fma_rate<T>, dim3(blocks), dim3(256), 0, nullptr, d, (T)0.999, (T)1e-4, iters);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (d) hipFree(d);
  if (e != hipSuccess) return fail(PU_ERR_HIP, std::string("pu_measure_fma_rate: ") + hipGetErrorString(e));
  *tflops = 2.0 * FMA_CHAINS * (double)iters * blocks * 256 * reps / (ms * 1e-3) / 1e12;
  return PU_OK;
}

}  // namespace pu

extern "C" {

int pu_measure_fma_rate(int precision, double *tflops) {
  if (!tflops) return pu::fail(PU_ERR_ARG, "null pointer");
  if (precision == PU_FP64) return pu::fma_rate<double>(tflops);
  if (precision == PU_FP32) return pu::fma_rate<float>(tflops);
  return pu::fail(PU_ERR_ARG, "precision must be PU_FP64 or PU_FP32");
}

int pu_measure_copy_rate(size_t bytes, double *gbps) {
  if (!gbps) return pu::fail(PU_ERR_ARG, "null pointer");
  if (bytes < (1u << 20) || bytes % 16) return pu::fail(PU_ERR_ARG, "bytes must be a multiple of 16, at least 1 MiB");
  void *a = nullptr, *b = nullptr;
  PU_HIP_CHECK(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return pu::fail(PU_ERR_HIP, "hipMalloc of the copy target failed");
  }
  int rc = pu::copy_device(b, a, bytes);   // warm-up (and first touch), waited for
  const int reps = 10;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = pu::fail(PU_ERR_HIP, "hipEventCreate failed");
  if (!rc && hipEventRecord(e0, nullptr) != hipSuccess) rc = pu::fail(PU_ERR_HIP, "hipEventRecord failed");
  for (int r = 0; r < reps && !rc; r++) rc = pu::launch_copy(b, a, bytes, nullptr);
  if (!rc && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = pu::fail(PU_ERR_HIP, "timing the copies failed");
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (rc) return rc;
  *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;     // bytes read + bytes written
  return PU_OK;
}

}  // extern "C"
