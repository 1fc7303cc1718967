// copy_probe.hip -- a plain 1 GiB device-to-device copy kernel, run three
// times.  Its FETCH_SIZE / WRITE_SIZE counter readings against the known
// 1 GiB per dispatch give the scale factors that turn those counters into HBM
// bytes for the advect kernels' traffic figures.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int main() {
  const size_t bytes = 1ull << 30, n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    std::printf("allocation failed\n");
    return 1;
  }
  if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("copied %zu bytes x3\n", bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
