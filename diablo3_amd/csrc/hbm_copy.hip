// hbm_copy.hip — plain device-to-device copy of a known byte count.
//
// Calibrates the rocprofv3 FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch)
// against a kernel whose HBM traffic is exactly 2 x bytes (read + write), and
// reports the copy bandwidth the device reaches.  Usage: hbm_copy [MiB] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const double4* __restrict__ src,
                                                   double4* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

#define CHK(x)                                                             \
  do {                                                                     \
    hipError_t e = (x);                                                    \
    if (e != hipSuccess) {                                                 \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));          \
      return 1;                                                            \
    }                                                                      \
  } while (0)

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 10;
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  CHK(hipMalloc(&a, bytes));
  CHK(hipMalloc(&b, bytes));
  CHK(hipMemset(a, 0, bytes));
  const int blocks = 256 * 32;
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  copy_kernel<<<blocks, 256>>>(a, b, n);
  CHK(hipDeviceSynchronize());
  CHK(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r) copy_kernel<<<blocks, 256>>>(a, b, n);
  CHK(hipEventRecord(e1));
  CHK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHK(hipEventElapsedTime(&ms, e0, e1));
  std::printf("hbm_copy: %zu bytes copied per dispatch (%zu read + %zu written), %d+1 dispatches, "
              "%.4f ms each, %.1f GB/s\n",
              bytes, bytes, bytes, reps, ms / reps, 2.0 * bytes * reps / (ms * 1e-3) / 1e9);
  CHK(hipFree(a));
  CHK(hipFree(b));
  return 0;
}
