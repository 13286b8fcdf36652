// hbm_copy — plain device-to-device copy of a known size, for calibrating the
// FETCH_SIZE / WRITE_SIZE HBM counters and measuring the copy bandwidth.
//
//   hbm_copy [MiB=256]
//
// Copies MiB MiB (float4 grid-stride loop) once untimed and 3 times timed with
// HIP events; prints {"bytes_read", "bytes_written", "ms", "GB/s"} as one JSON
// line.  Each timed dispatch reads and writes MiB MiB exactly once.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
      return 1;                                                              \
    }                                                                        \
  } while (0)

__global__ void __launch_bounds__(256) k_copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill(float4* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_float4((float)i, 1.0f, 2.0f, 3.0f);
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 256;
  if (mib == 0 || mib > 65536) {
    std::fprintf(stderr, "size must be 1..65536 MiB\n");
    return 2;
  }
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int grid = prop.multiProcessorCount * 8;
  hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, 0, a, n);
  hipLaunchKernelGGL(k_copy_f4, dim3(grid), dim3(256), 0, 0, a, b, n);
  CHECK(hipGetLastError());
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_copy_f4, dim3(grid), dim3(256), 0, 0, a, b, n);
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  ms /= 3;
  // spot check the copy
  float4 h[2];
  CHECK(hipMemcpy(h, b + (n - 2), sizeof h, hipMemcpyDeviceToHost));
  if (h[1].x != (float)(n - 1) || h[1].w != 3.0f) {
    std::fprintf(stderr, "copy mismatch\n");
    return 1;
  }
  std::printf("{\"kernel\": \"k_copy_f4\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"ms\": %.6f, \"GB/s\": %.1f}\n",
              bytes, bytes, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
