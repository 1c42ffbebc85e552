// Calibration program for HBM traffic counters: a plain float4 copy of 1 GiB
// (one float4 per thread, 2^16 blocks of 256), 3 warm-up and 20 timed
// launches.  Under
// `rocprofv3 --pmc` the request-size counters of each dispatch must read back
// 1 GiB read and 1 GiB written; the printed rate is the copy bandwidth of the
// device (read + write bytes per second).
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void __launch_bounds__(256) copyk(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

#define CK(call)                                                                   \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));       \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

int main() {
  const size_t bytes = 1ull << 30;
  const size_t n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  for (int it = 0; it < 3; ++it) hipLaunchKernelGGL(copyk, dim3((unsigned)(n / 256)), dim3(256), 0, 0, a, b, n);
  CK(hipGetLastError());
  const int reps = 20;
  CK(hipEventRecord(e0));
  for (int it = 0; it < reps; ++it) hipLaunchKernelGGL(copyk, dim3((unsigned)(n / 256)), dim3(256), 0, 0, a, b, n);
  CK(hipEventRecord(e1));
  CK(hipEventSynchronize(e1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, e0, e1));
  std::printf("copy 1 GiB: %.3f ms per launch, %.1f GB/s (read + write)\n", ms / reps,
              2.0 * bytes / (ms / reps * 1e-3) / 1e9);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
