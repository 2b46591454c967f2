// probe.hip — measured ceilings of the GPU the library runs on, reported next
// to the published MI355X peaks: sustained FP64 / FP32 VALU FMA rate (many
// independent chains per lane, no memory traffic) and a plain device copy.
#include <hip/hip_runtime.h>

#include "common.h"

namespace fold3d {

constexpr int kChains = 16;
constexpr int kIters = 2048;

template <typename R>
__global__ __launch_bounds__(256) void fma_probe_kernel(R *out, R a, R b) {
  R acc[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) acc[c] = R(threadIdx.x + c) * R(1e-3);
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc[c] = fma(acc[c], a, b);
  }
  R s = R(0);
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += acc[c];
  if (s == R(-1)) out[blockIdx.x * blockDim.x + threadIdx.x] = s;  // keeps the chains live
}

__global__ __launch_bounds__(256) void copy_probe_kernel(const double4 *__restrict__ in,
                                                         double4 *__restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

template <typename R>
static int time_fma(hipStream_t st, double *flops) {
  int dev = 0, cus = 0;
  F3D_HIP(hipGetDevice(&dev));
  F3D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = cus * 32;  // 8 waves per SIMD
  R *out = nullptr;
  F3D_HIP(hipMalloc(&out, sizeof(R) * blocks * 256));
  hipEvent_t e0, e1;
  F3D_HIP(hipEventCreate(&e0));
  F3D_HIP(hipEventCreate(&e1));
  float ms = 0;
  for (int rep = 0; rep < 2; ++rep) {
    F3D_HIP(hipEventRecord(e0, st));
    fma_probe_kernel<R><<<blocks, 256, 0, st>>>(out, R(0.999999), R(1e-7));
    F3D_HIP(hipEventRecord(e1, st));
    F3D_HIP(hipEventSynchronize(e1));
    F3D_HIP(hipEventElapsedTime(&ms, e0, e1));
  }
  *flops = 2.0 * kChains * (double)kIters * blocks * 256 / (ms * 1e-3);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  F3D_HIP(hipFree(out));
  return FOLD3D_OK;
}

}  // namespace fold3d

using namespace fold3d;

extern "C" int fold3d_measure_peaks(void *stream, double *fp64_flops, double *fp32_flops,
                                    double *copy_bytes_per_s) {
  hipStream_t st = (hipStream_t)stream;
  int rc = FOLD3D_OK;
  if (fp64_flops && (rc = time_fma<double>(st, fp64_flops)) != FOLD3D_OK) return rc;
  if (fp32_flops && (rc = time_fma<float>(st, fp32_flops)) != FOLD3D_OK) return rc;
  if (copy_bytes_per_s) {
    const size_t bytes = size_t(2) << 30;  // 2 GiB each way
    void *a = nullptr, *b = nullptr;
    F3D_HIP(hipMalloc(&a, bytes));
    F3D_HIP(hipMalloc(&b, bytes));
    F3D_HIP(hipMemsetAsync(a, 0, bytes, st));
    hipEvent_t e0, e1;
    F3D_HIP(hipEventCreate(&e0));
    F3D_HIP(hipEventCreate(&e1));
    int dev = 0, cus = 0;
    F3D_HIP(hipGetDevice(&dev));
    F3D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    float ms = 0;
    for (int rep = 0; rep < 3; ++rep) {
      F3D_HIP(hipEventRecord(e0, st));
      copy_probe_kernel<<<cus * 16, 256, 0, st>>>((const double4 *)a, (double4 *)b,
                                                   bytes / sizeof(double4));
      F3D_HIP(hipEventRecord(e1, st));
      F3D_HIP(hipEventSynchronize(e1));
      F3D_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    *copy_bytes_per_s = 2.0 * bytes / (ms * 1e-3);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    F3D_HIP(hipFree(a));
    F3D_HIP(hipFree(b));
  }
  return FOLD3D_OK;
}
