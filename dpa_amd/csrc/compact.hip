// compact.hip -- launchers of the K_compact kernels (compact_kernels.hpp).
#include "compact_kernels.hpp"
#include "runtime.hpp"

namespace dpah {

Chunk compact_rows(dpah_context* ctx, const float* slots, uint64_t stride, const uint32_t* steps,
                   const uint8_t* special, uint64_t count, uint64_t vertex_base, int use64,
                   hipStream_t stream) {
  hipStream_t st = stream ? stream : ctx->stream;
  auto sync = [&] {
    DPAH_HIP(hipStreamSynchronize(st));
    ctx->resolve_timings();
  };
  Chunk ch;
  if (count == 0) return ch;
  DeviceBuffer vcount(ctx, count * sizeof(uint64_t));
  DeviceBuffer icount(ctx, count * sizeof(uint64_t));
  const uint64_t blocks = (count + kScanTile - 1) / kScanTile;
  DeviceBuffer sums(ctx, (2 * blocks + 2) * sizeof(uint64_t));

  CompactArgs a{};
  a.slots = slots;
  a.stride = stride;
  a.steps = steps;
  a.special = special;
  a.count = count;
  a.vcount = vcount.as<uint64_t>();
  a.icount = icount.as<uint64_t>();
  a.use64 = use64;
  a.vertex_base = vertex_base;

  const unsigned g1 = (unsigned)((count + 255) / 256);
  hipLaunchKernelGGL(k_compact_count, dim3(g1), dim3(256), 0, st, a);
  DPAH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)blocks), dim3(kScanThreads), 0, st,
                     a.vcount, a.icount, count, sums.as<uint64_t>());
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, sums.as<uint64_t>(),
                     blocks);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)blocks), dim3(kScanThreads), 0, st,
                     a.vcount, a.icount, count, sums.as<uint64_t>());
  DPAH_HIP(hipGetLastError());
  uint64_t totals[2] = {0, 0};
  DPAH_HIP(hipMemcpyAsync(totals, sums.as<uint64_t>() + 2 * blocks, sizeof(totals),
                          hipMemcpyDeviceToHost, st));
  sync();
  ch.V = totals[0];
  ch.I = 2 * totals[1];
  const uint64_t ib = use64 ? 8 : 4;
  ch.bytes = ch.V * 12 + ch.V * 4 + ch.I * ib;
  // one allocation per chunk, carved into vertices | colors | indices
  if (ch.bytes) {
    void* base = ctx->alloc(ch.bytes);
    ch.vertices = static_cast<float*>(base);
    ch.colors = ch.vertices + 3 * ch.V;
    ch.indices = ch.colors + ch.V;
  }
  a.vertices = ch.vertices;
  a.colors = ch.colors;
  a.indices = ch.indices;
  // one workgroup per group of 8 rows of a 64-row tile and span of slots (a
  // row has at most `stride` slots); tiles padded to a multiple of the XCD count
  const uint64_t tiles = (count + kTileRows - 1) / kTileRows;
  const uint64_t spans = (stride + kCompactSpan - 1) / kCompactSpan;
  const uint64_t xcds = (uint64_t)std::max(1, ctx->xcds);
  const uint64_t groups = (tiles + xcds - 1) / xcds * xcds * spans * kGroupsPerTile;
  DPAH_REQUIRE(groups < (1ull << 31), "too many rows or slots in one compaction");
  a.tiles = (uint32_t)tiles;
  a.spans = (uint32_t)spans;
  a.xcds = (uint32_t)xcds;
  ctx->timed("compact_write", st, [&] {
    hipLaunchKernelGGL(k_compact_write, dim3((unsigned)groups), dim3(kCompactThreads), 0, st, a);
  });
  DPAH_HIP(hipGetLastError());
  sync();
  return ch;
}

}  // namespace dpah
