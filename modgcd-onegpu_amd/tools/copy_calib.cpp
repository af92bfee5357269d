// copy_calib -- a plain device copy of known size, for calibrating the
// FETCH_SIZE / WRITE_SIZE counters of rocprofv3: each of the three launches
// reads 256 MiB and writes 256 MiB (262,144 KiB each way).
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void fill_kernel(uint4* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(v, v, v, v);
}

int main() {
  const size_t bytes = 256ull << 20, n = bytes / sizeof(uint4);
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, 0, a, n, 1u);
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, 0, b, n, 0u);
  for (int i = 0; i < 3; ++i) hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("copy_kernel x3: %zu KiB read and %zu KiB written per launch\n", bytes >> 10, bytes >> 10);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
