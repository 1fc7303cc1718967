// compact.hip -- launchers of the K_compact kernels (compact_kernels.hpp).
#include <algorithm>

#include "compact_kernels.hpp"
#include "runtime.hpp"

namespace dpah {

uint64_t compact_sums_words(uint64_t count) {
  return 2 * ((count + kScanTile - 1) / kScanTile) + 2;
}

void compact_count(dpah_context* ctx, const CompactRows& r, uint64_t* vcount, uint64_t* icount,
                   uint64_t* sums, hipStream_t st) {
  if (r.count == 0) return;
  CompactArgs a{};
  a.slots = r.slots;
  a.stride = r.stride;
  a.steps = r.steps;
  a.special = r.special;
  a.count = r.count;
  a.vcount = vcount;
  a.icount = icount;
  a.use64 = r.use64;
  a.vertex_base = r.vertex_base;
  a.slot_shift = r.slot_shift;
  const uint64_t blocks = (r.count + kScanTile - 1) / kScanTile;
  const unsigned g1 = (unsigned)((r.count + 255) / 256);
  hipLaunchKernelGGL(k_compact_count, dim3(g1), dim3(256), 0, st, a);
  DPAH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)blocks), dim3(kScanThreads), 0, st, vcount,
                     icount, r.count, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, sums, blocks);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)blocks), dim3(kScanThreads), 0, st, vcount,
                     icount, r.count, sums);
  DPAH_HIP(hipGetLastError());
  (void)ctx;
}

namespace {

// The chunk (one allocation: vertices | colours | indices) and the launch
// arguments of one round's writes; returns the workgroup count, 0 for an
// empty round.
uint64_t plan_write(dpah_context* ctx, const CompactRows& r, const uint64_t* vcount,
                    const uint64_t* icount, uint64_t vertices, uint64_t segments, Chunk& ch,
                    CompactArgs& a) {
  ch = Chunk{};
  a = CompactArgs{};
  if (r.count == 0) return 0;
  ch.V = vertices;
  ch.I = 2 * segments;
  const uint64_t ib = r.use64 ? 8 : 4;
  ch.bytes = ch.V * 12 + ch.V * 4 + ch.I * ib;
  if (ch.bytes) {
    void* base = ctx->alloc(ch.bytes);
    ch.vertices = static_cast<float*>(base);
    ch.colors = ch.vertices + 3 * ch.V;
    ch.indices = ch.colors + ch.V;
  }
  a.slots = r.slots;
  a.stride = r.stride;
  a.steps = r.steps;
  a.special = r.special;
  a.count = r.count;
  a.vcount = const_cast<uint64_t*>(vcount);
  a.icount = const_cast<uint64_t*>(icount);
  a.slot_shift = r.slot_shift;
  a.use64 = r.use64;
  a.vertex_base = r.vertex_base;
  a.vertices = ch.vertices;
  a.colors = ch.colors;
  a.indices = ch.indices;
  // one workgroup per group of rows of a 64-row tile and span of slots (a
  // row has at most `stride` slots); tiles padded to a multiple of the XCD
  // count.  Spans are whole rows when the row groups alone give 32 workgroups
  // per CU; fewer groups split their rows into spans of whole chunks, at most
  // as many as make up that count (a span past all its rows' ends costs a
  // workgroup that only reads the rows' lengths)
  const uint64_t tiles = (r.count + kTileRows - 1) / kTileRows;
  const uint64_t xcds = (uint64_t)std::max(1, ctx->xcds);
  const uint64_t row_groups = (tiles + xcds - 1) / xcds * xcds * kGroupsPerTile;
  const uint64_t want = (uint64_t)std::max(1, ctx->cu_count) * 32;
  const uint64_t chunks = (r.stride + kCompactChunk - 1) / kCompactChunk;
  const uint64_t split = std::min(chunks, (want + row_groups - 1) / row_groups);
  const uint64_t span_slots = (chunks + split - 1) / split * kCompactChunk;
  const uint64_t spans = (r.stride + span_slots - 1) / span_slots;
  const uint64_t groups = row_groups * spans;   // a multiple of xcds
  DPAH_REQUIRE(span_slots < (1ull << 31), "rows too long for one compaction");
  DPAH_REQUIRE(groups < (1ull << 31), "too many rows or slots in one compaction");
  a.tiles = (uint32_t)tiles;
  a.spans = (uint32_t)spans;
  a.span_slots = (uint32_t)span_slots;
  a.xcds = (uint32_t)xcds;
  return groups;
}

}  // namespace

Chunk compact_write(dpah_context* ctx, const CompactRows& r, const uint64_t* vcount,
                    const uint64_t* icount, uint64_t vertices, uint64_t segments,
                    hipStream_t st) {
  Chunk ch;
  CompactArgs a;
  const uint64_t groups = plan_write(ctx, r, vcount, icount, vertices, segments, ch, a);
  if (!groups) return ch;
  ctx->timed("compact_write", st, [&] {
    hipLaunchKernelGGL(k_compact_write, dim3((unsigned)groups), dim3(kCompactThreads), 0, st, a);
  });
  DPAH_HIP(hipGetLastError());
  return ch;
}

std::vector<Chunk> compact_write_many(dpah_context* ctx, const std::vector<WriteJob>& jobs,
                                      DeviceBuffer& table, hipStream_t st) {
  std::vector<Chunk> out(jobs.size());
  std::vector<CompactSeg> segs;
  uint64_t first = 0;
  try {
    for (size_t i = 0; i < jobs.size(); ++i) {
      CompactSeg g{};
      const uint64_t groups = plan_write(ctx, jobs[i].rows, jobs[i].vcount, jobs[i].icount,
                                         jobs[i].vertices, jobs[i].segments, out[i], g.a);
      if (!groups) continue;
      g.first = (uint32_t)first;
      first += groups;
      DPAH_REQUIRE(first < (1ull << 31), "too many workgroups in one compaction launch");
      segs.push_back(g);
    }
  } catch (...) {
    for (Chunk& c : out)
      if (c.vertices) ctx->release(c.vertices, c.bytes);
    throw;
  }
  if (segs.empty()) return out;
  try {
    table = DeviceBuffer(ctx, segs.size() * sizeof(CompactSeg));
    DPAH_HIP(hipMemcpyAsync(table.ptr, segs.data(), segs.size() * sizeof(CompactSeg),
                            hipMemcpyHostToDevice, st));
    const CompactSeg* d = table.as<CompactSeg>();
    const uint32_t n = (uint32_t)segs.size();
    ctx->timed("compact_write", st, [&] {
      hipLaunchKernelGGL(k_compact_write_many, dim3((unsigned)first), dim3(kCompactThreads), 0,
                         st, d, n);
    });
    DPAH_HIP(hipGetLastError());
  } catch (...) {
    for (Chunk& c : out)
      if (c.vertices) ctx->release(c.vertices, c.bytes);
    throw;
  }
  return out;
}

Chunk compact_rows(dpah_context* ctx, const float* slots, uint64_t stride, uint32_t g,
                   const uint32_t* steps, const uint8_t* special, uint64_t count,
                   uint64_t vertex_base, int use64, hipStream_t stream) {
  hipStream_t st = stream ? stream : ctx->stream;
  auto sync = [&] {
    DPAH_HIP(hipStreamSynchronize(st));
    ctx->resolve_timings();
  };
  if (count == 0) return Chunk{};
  const CompactRows r{slots, stride, steps, special, count, vertex_base, use64, g};
  DeviceBuffer vcount(ctx, count * sizeof(uint64_t));
  DeviceBuffer icount(ctx, count * sizeof(uint64_t));
  const uint64_t words = compact_sums_words(count);
  DeviceBuffer sums(ctx, words * sizeof(uint64_t));
  compact_count(ctx, r, vcount.as<uint64_t>(), icount.as<uint64_t>(), sums.as<uint64_t>(), st);
  uint64_t totals[2] = {0, 0};   // the scan leaves them in the last two words
  DPAH_HIP(hipMemcpyAsync(totals, sums.as<uint64_t>() + words - 2, sizeof(totals),
                          hipMemcpyDeviceToHost, st));
  sync();
  Chunk ch = compact_write(ctx, r, vcount.as<uint64_t>(), icount.as<uint64_t>(), totals[0],
                           totals[1], st);
  sync();
  return ch;
}

}  // namespace dpah
