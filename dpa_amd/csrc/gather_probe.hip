// gather_probe.hip -- the access pattern of a cell-record reload on its own:
// each lane reads six float4 (96 B, one cell's eight corners) from randomly
// chosen records of a 2 GiB array, 16 records per lane, 2^20 lanes.  Two
// record strides: 128 B (the cell layout's padded records, one aligned line
// each) and 96 B (packed records, which straddle two lines half the time).
// Each stride runs three times; the time per dispatch and its FETCH_SIZE
// counter give the rate and the bytes that random whole-line reads reach on
// this chip, the practical ceiling of the advect kernel's cell reloads.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr uint32_t kRecords = 1u << 24;   // 2 GiB at 128 B per record
constexpr uint32_t kLanes = 1u << 20;
constexpr int kPerLane = 16;

__device__ inline uint32_t scramble(uint32_t x) {   // an invertible 32-bit integer mix
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

template <int kStride>
__global__ void k_gather(const float4* __restrict__ base, float* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  float acc = 0.0f;
  for (int j = 0; j < kPerLane; ++j) {
    const uint32_t r = scramble(t * 131u + (uint32_t)j * 0x9e3779b9u) & (kRecords - 1);
    const float4* p = base + (uint64_t)r * (kStride / 16);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], g = p[5];
    acc += a.x + a.y + a.z + a.w + b.x + b.y + b.z + b.w + c.x + c.y + c.z + c.w;
    acc += d.x + d.y + d.z + d.w + e.x + e.y + e.z + e.w + g.x + g.y + g.z + g.w;
  }
  out[t] = acc;
}

}  // namespace

int main() {
  const size_t bytes = (size_t)kRecords * 128;
  float4* base = nullptr;
  float* out = nullptr;
  if (hipMalloc(&base, bytes) != hipSuccess || hipMalloc(&out, kLanes * sizeof(float)) != hipSuccess) {
    std::printf("allocation failed\n");
    return 1;
  }
  if (hipMemset(base, 1, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  const double records = (double)kLanes * kPerLane;
  for (int rep = 0; rep < 3; ++rep) {
    for (int stride : {128, 96}) {
      (void)hipEventRecord(e0, 0);
      if (stride == 128)
        hipLaunchKernelGGL(k_gather<128>, dim3(kLanes / 256), dim3(256), 0, 0, base, out);
      else
        hipLaunchKernelGGL(k_gather<96>, dim3(kLanes / 256), dim3(256), 0, 0, base, out);
      if (hipGetLastError() != hipSuccess) return 1;
      (void)hipEventRecord(e1, 0);
      if (hipEventSynchronize(e1) != hipSuccess) return 1;
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      std::printf("stride %3d B: %.3f ms, %.3e records/s, %.1f GB/s of whole 128-B lines "
                  "at one line per record\n",
                  stride, ms, records / (ms * 1e-3), records * 128 / (ms * 1e6));
      std::fflush(stdout);
    }
  }
  (void)hipFree(base);
  (void)hipFree(out);
  return 0;
}
