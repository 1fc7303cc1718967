// compact_kernels.hpp -- K_compact: prune_integral_curves
// (particle_advector.cpp:402-408), index_generator::generate
// (index_generator.cpp:10-39) and color_generator::generate_from_angular_velocities
// (color_generator.cpp:37-58) fused over one chunk of a round's slots.
//
// Slot row p (stride s, the block layout of slot_float in advect_kernels.hpp)
// holds particle p's seed, one vertex per step and the terminal vertex at
// k_p + 1; rows are never pre-filled, so the kernels read only slots
// [0, k_p + 1].  For a row with no sentinel-valued vertex the pruned curve is
// the whole row, it contributes k_p + 2 vertices and k_p segments
// (b + j, b + j + 1), j < k_p.  A row holding a vertex equal to (-1,-1,-1) or
// (-2,-2,-2) -- dpa would drop / treat it as terminal -- is walked serially with
// the literal rules.
#pragma once

#include "advect_kernels.hpp"

namespace dpah {

struct CompactArgs {
  const float*    slots;     // slot layout of slot_float, whole tiles
  uint64_t        stride;
  const uint32_t* steps;     // k_p
  const uint8_t*  special;   // row holds a sentinel-valued vertex
  uint64_t        count;
  uint64_t*       vcount;    // per row: kept vertices   (scanned in place to offsets)
  uint64_t*       icount;    // per row: index pairs     (scanned in place to offsets)
  float*          vertices;  // [V][3] of this chunk
  float*          colors;    // [V]
  void*           indices;   // [2 I] u32 or u64
  int             use64;
  uint64_t        vertex_base;  // round-level index of this chunk's first vertex
  uint32_t        tiles;        // ceil(count / 64)
  uint32_t        spans;        // slot spans per tile
  uint32_t        span_slots;   // slots per span, a multiple of kCompactChunk
  uint32_t        xcds;         // workgroups go to the XCDs round-robin by id
  uint32_t        slot_shift;   // slot blocks of 2^slot_shift (slot_float)
};

__device__ __forceinline__ bool eq3(const float* v, float s) {
  return v[0] == s && v[1] == s && v[2] == s;
}

// Eigen's fixed-size-3 reductions are x0 + (x1 + x2).
__device__ __forceinline__ float sum3(float a, float b, float c) {
  const float bc = b + c;
  return a + bc;
}

__device__ __forceinline__ void normalized3(float x, float y, float z, float o[3]) {
  const float xx = x * x, yy = y * y, zz = z * z;
  const float q = sum3(xx, yy, zz);
  if (q > 0.0f) {
    const float s = sqrtf(q);
    o[0] = x / s; o[1] = y / s; o[2] = z / s;
  } else {
    o[0] = x; o[1] = y; o[2] = z;
  }
}

// atan2(|prev x next|, prev . next) with the Eigen cross / dot / norm order.
__device__ __forceinline__ float angle_color(const float prev[3], const float next[3]) {
  float c[3], p, q;
  p = prev[1] * next[2]; q = prev[2] * next[1]; c[0] = p - q;
  p = prev[2] * next[0]; q = prev[0] * next[2]; c[1] = p - q;
  p = prev[0] * next[1]; q = prev[1] * next[0]; c[2] = p - q;
  const float c0 = c[0] * c[0], c1 = c[1] * c[1], c2 = c[2] * c[2];
  const float nrm = sqrtf(sum3(c0, c1, c2));
  const float d0 = prev[0] * next[0], d1 = prev[1] * next[1], d2 = prev[2] * next[2];
  return atan2f(nrm, sum3(d0, d1, d2));
}

__device__ __forceinline__ void put_index(const CompactArgs& a, uint64_t pos, uint64_t g) {
  if (a.use64) {
    uint64_t* o = (uint64_t*)a.indices;
    o[2 * pos] = g; o[2 * pos + 1] = g + 1;
  } else {
    uint32_t* o = (uint32_t*)a.indices;
    o[2 * pos] = (uint32_t)g; o[2 * pos + 1] = (uint32_t)(g + 1);
  }
}

// Counts for one row (thread per row).
static __global__ __launch_bounds__(256) void k_compact_count(CompactArgs a) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.count) return;
  const uint64_t k = a.steps[p];
  if (!a.special[p]) {
    a.vcount[p] = k + 2;
    a.icount[p] = k;
    return;
  }
  const float* row = a.slots + slot_float(p, a.stride, 0, a.slot_shift);
  uint64_t kept = 0, pairs = 0;
  bool have_prev = false, prev_term = false;
  for (uint64_t j = 0; j <= k + 1; ++j) {
    const float* v = row + slot_offset(j, a.slot_shift);
    if (eq3(v, -1.0f)) continue;
    const bool term = eq3(v, -2.0f);
    if (have_prev && !prev_term && !term) ++pairs;
    have_prev = true;
    prev_term = term;
    ++kept;
  }
  a.vcount[p] = kept;
  a.icount[p] = pairs;
}

// Serial walk of a row that holds sentinel-valued vertices: literal prune,
// index and colour rules on the row's kept vertices.
static __device__ void compact_special_row(const CompactArgs& a, uint64_t p) {
  const uint64_t k = a.steps[p];
  const float* row = a.slots + slot_float(p, a.stride, 0, a.slot_shift);
  const uint64_t vb = a.vcount[p], ib = a.icount[p];
  // list of kept slot indices is walked twice with a small lookahead
  uint64_t m = 0, pairs = 0;
  int64_t prev_j = -1;
  for (uint64_t j = 0; j <= k + 1; ++j) {
    const float* v = row + slot_offset(j, a.slot_shift);
    if (eq3(v, -1.0f)) continue;
    // next kept slot
    int64_t next_j = -1;
    for (uint64_t q = j + 1; q <= k + 1; ++q)
      if (!eq3(row + slot_offset(q, a.slot_shift), -1.0f)) { next_j = (int64_t)q; break; }
    const uint64_t g = a.vertex_base + vb + m;
    float* out = a.vertices + 3 * (vb + m);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    float col = 0.0f;
    const bool term = eq3(v, -2.0f);
    if (!term) {
      float prev[3] = {0.0f, 0.0f, 0.0f}, next[3] = {0.0f, 0.0f, 0.0f};
      if (g > 1 && prev_j >= 0 && !eq3(row + slot_offset((uint64_t)prev_j, a.slot_shift), -2.0f)) {
        const float* u = row + slot_offset((uint64_t)prev_j, a.slot_shift);
        normalized3(v[0] - u[0], v[1] - u[1], v[2] - u[2], prev);
      }
      if (next_j >= 0 && !eq3(row + slot_offset((uint64_t)next_j, a.slot_shift), -2.0f)) {
        const float* u = row + slot_offset((uint64_t)next_j, a.slot_shift);
        normalized3(u[0] - v[0], u[1] - v[1], u[2] - v[2], next);
        put_index(a, ib + pairs, g);
        ++pairs;
      }
      col = angle_color(prev, next);
    }
    a.colors[vb + m] = col;
    prev_j = (int64_t)j;
    ++m;
  }
}

// Writes: one workgroup (four waves) per group of kGroupRows = 4 rows of a
// 64-row tile and span of span_slots slots (whole rows when the launch has
// enough groups to fill the chip, so that no workgroup starts on slots past
// its rows' ends), walked in chunks of kCompactChunk = 63 vertices; wave h
// owns row h of the group.  Few rows per
// workgroup keep few output runs open at once (3 streams per row): on
// 2^18 x 1000 recorded rows, 64 rows per workgroup left so many lines partly
// written that they outgrew L2 (stores at 2.6 TB/s, compaction 2.2 ms); 8
// rows: 1.74 ms; 4 rows: 1.66 ms; 2 rows: 1.83 ms.  Workgroup ids map so that
// the groups of a span run together, and the spans of a tile one after
// another, on one XCD: the groups share the slot lines of their span in L2,
// and the lines that two spans of a row share are completed in one L2.
//
// A chunk's slots [j0 - 1, j0 + kCompactChunk] of the group's rows are staged
// in LDS row by row (row pitch kBandRow floats, odd, so lanes reading
// consecutive slots of one row hit different banks).  Lane i computes the
// direction d = normalized(x[j0+i] - x[j0+i-1]), which is the
// previous-segment direction of vertex j0 + i and, fetched from lane i + 1,
// the next-segment direction of vertex j0 + i - 1 (the same expression on
// the same operands, so each direction is computed once).  Lanes 0..62 emit
// vertices j0 .. j0 + 62.  Rows with sentinel-valued vertices are walked
// serially by the span-0 workgroup of their group, one thread each.
constexpr uint32_t kRunLanes = 64;                     // lanes per row (a wave)
constexpr uint32_t kCompactThreads = 256;              // workgroup size
constexpr uint32_t kGroupRows = kCompactThreads / kRunLanes;   // rows per workgroup
constexpr uint32_t kCompactChunk = kRunLanes - 1;
constexpr uint32_t kGroupsPerTile = kTileRows / kGroupRows;
constexpr uint32_t kBandRow = 3 * (kCompactChunk + 2);   // staged slots x 3 floats, odd
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));

// The body of one write workgroup; `block` is its workgroup index within the
// launch of `a` (blockIdx.x for a launch of its own).
__device__ __forceinline__ void compact_write_block(const CompactArgs& a, uint32_t block) {
  __shared__ float buf[kGroupRows * kBandRow];
  __shared__ uint32_t len[kGroupRows];   // k + 2 slots to emit, 0: none (absent or special)
  __shared__ uint64_t vbase[kGroupRows], ibase[kGroupRows];
  const uint32_t t = threadIdx.x;
  const uint32_t xcd = block % a.xcds, local = block / a.xcds;
  const uint32_t per_tile = a.spans * kGroupsPerTile;
  const uint32_t tile = (local / per_tile) * a.xcds + xcd, rest = local % per_tile;
  const uint32_t span = rest / kGroupsPerTile, grp = rest % kGroupsPerTile;
  if (tile >= a.tiles) return;   // padding of the last tiles (whole workgroup)
  const uint64_t row0 = (uint64_t)tile * kTileRows + grp * kGroupRows;
  if (t < kGroupRows) {
    const uint64_t r = row0 + t;
    const bool live = r < a.count && !a.special[r];
    len[t] = live ? a.steps[r] + 2 : 0u;
    vbase[t] = live ? a.vcount[r] : 0;
    ibase[t] = live ? a.icount[r] : 0;
  }
  __syncthreads();
  uint32_t nmax = 0;
  for (uint32_t i = 0; i < kGroupRows; ++i) nmax = max(nmax, len[i]);
  nmax = (uint32_t)__builtin_amdgcn_readfirstlane(nmax);   // same in every lane
  // slot j of group row l: tile_slot_float(j, b0, grp * kGroupRows + l, gs) past the
  // tile's block b0, gs = a.slot_shift
  const uint32_t gs = a.slot_shift;
  const float* tile_base = a.slots + slot_float((uint64_t)tile * kTileRows, a.stride, 0, gs);
  const uint32_t h = t / kRunLanes, i = t % kRunLanes;
  const uint32_t jb = span * a.span_slots;   // < stride
  const uint32_t je = (uint32_t)min((uint64_t)nmax, (uint64_t)jb + a.span_slots);

  // Staging: item x = t, t + kCompactThreads of the chunk's (slot, row)
  // pairs, slot lo + x / kGroupRows of row x % kGroupRows, for slots
  // [lo, hi] = [j0 - 1, j0 + kCompactChunk]
  // clamped to the rows' extent.  Only slots below the row's length are read:
  // they are all a row ever uses.  The loads are buffer loads over the
  // chunk's slots; an item past its row's end gets an offset beyond the
  // buffer's size, which the hardware answers with zeros without touching
  // memory.  A chunk's loads are issued while the previous chunk computes
  // (software pipeline); slot s goes to staged slot s - (j0 - 1).
  constexpr uint32_t kPer = ((kCompactChunk + 2) * kGroupRows + kCompactThreads - 1) / kCompactThreads;
  u32x3 r[kPer];
  auto chunk_lo = [](uint32_t j0) { return j0 == 0 ? 0u : j0 - 1; };
  auto chunk_hi = [nmax](uint32_t j0) { return min(j0 + kCompactChunk, nmax - 1); };
  auto issue = [&](uint32_t j0) {
    const uint32_t lo = chunk_lo(j0), hi = chunk_hi(j0);
    // base and size are the same in every lane; readfirstlane says so, so the
    // resource sits in scalar registers (readfirstlane returns int: each half
    // goes through uint32_t before widening, or bit 31 of the low half would
    // be sign-extended into the high half)
    const uint32_t b0 = lo >> gs;   // the chunk's blocks b0 .. hi >> gs
    const uint32_t nbytes = (uint32_t)__builtin_amdgcn_readfirstlane(
        ((hi >> gs) - b0 + 1) * ((uint32_t)kSlotPitch << gs) * 4u);   // <= 65 * 768 or 10 * 6144
    const uint64_t base = (uint64_t)(tile_base + (uint64_t)b0 * (kSlotPitch << gs));
    const uint32_t base_lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)base);
    const uint32_t base_hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
    const uint64_t ubase = ((uint64_t)base_hi << 32) | (uint64_t)base_lo;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)ubase, (short)0, (int)nbytes, kBufferWord3);
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const uint32_t x = t + kCompactThreads * q, l = x % kGroupRows, s = lo + x / kGroupRows;
      const bool ok = s <= hi && s < len[l];
      const uint32_t off = ok ? tile_slot_float(s, b0, grp * kGroupRows + l, gs) * 4u : 0x80000000u;
      r[q] = __builtin_amdgcn_raw_buffer_load_b96(rs, (int)off, 0, 0);
    }
  };
  if (jb < je) issue(jb);
  for (uint32_t j0 = jb; j0 < je; j0 += kCompactChunk) {
    {
      const uint32_t lo = chunk_lo(j0), hi = chunk_hi(j0);
#pragma unroll
      for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t x = t + kCompactThreads * q, l = x % kGroupRows, s = lo + x / kGroupRows;
        if (s <= hi && s < len[l]) {
          float* d = buf + l * kBandRow + 3 * (s + 1 - j0);
          d[0] = __uint_as_float(r[q].x); d[1] = __uint_as_float(r[q].y);
          d[2] = __uint_as_float(r[q].z);
        }
      }
    }
    __syncthreads();
    if (j0 + kCompactChunk < je) issue(j0 + kCompactChunk);   // in flight during this chunk
    const uint32_t n = len[h];               // vertices j < n; colours and directions j <= n - 2
    if (n > j0) {                            // uniform over the row's lanes
      const uint32_t j = j0 + i;             // this lane's vertex
      const float* v = buf + h * kBandRow + 3 * (i + 1);
      float d[3] = {0.0f, 0.0f, 0.0f};
      if (j >= 1 && j + 2 <= n) {            // 1 <= j <= k
        const float* u = v - 3;
        normalized3(v[0] - u[0], v[1] - u[1], v[2] - u[2], d);
      }
      const float n0 = __shfl_down(d[0], 1, kRunLanes), n1 = __shfl_down(d[1], 1, kRunLanes),
                  n2 = __shfl_down(d[2], 1, kRunLanes);
      if (i < kCompactChunk && j < n) {
        const uint64_t vb = vbase[h], ib = ibase[h];
        const float x = v[0], y = v[1], z = v[2];
        float* out = a.vertices + 3 * (vb + j);
        out[0] = x; out[1] = y; out[2] = z;
        float col = 0.0f;
        if (j + 2 <= n) {                    // j <= k
          const uint64_t g = a.vertex_base + vb + j;
          float prev[3] = {0.0f, 0.0f, 0.0f}, next[3] = {0.0f, 0.0f, 0.0f};
          if (j >= 1 && g > 1) { prev[0] = d[0]; prev[1] = d[1]; prev[2] = d[2]; }
          if (j + 3 <= n) {                  // j + 1 <= k
            next[0] = n0; next[1] = n1; next[2] = n2;
            put_index(a, ib + j, g);
          }
          col = angle_color(prev, next);
        }
        a.colors[vb + j] = col;
      }
    }
    __syncthreads();
  }
  if (span == 0 && t < kGroupRows) {
    const uint64_t rr = row0 + t;
    if (rr < a.count && a.special[rr]) compact_special_row(a, rr);
  }
}

static __global__ __launch_bounds__(kCompactThreads) void k_compact_write(CompactArgs a) {
  compact_write_block(a, blockIdx.x);
}

// The writes of several rounds in one launch: segment s owns workgroups
// [first, first of s + 1); each segment's workgroup count is a multiple of
// the XCD count, so block % xcds is the same XCD as in a launch of its own,
// and every segment's workgroups do exactly what they would do there.
struct CompactSeg {
  CompactArgs a;
  uint32_t first;
};

static __global__ __launch_bounds__(kCompactThreads) void k_compact_write_many(
    const CompactSeg* segs, uint32_t nseg) {
  uint32_t lo = 0, hi = nseg;   // last segment with first <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].first <= blockIdx.x) lo = mid; else hi = mid;
  }
  const CompactArgs a = segs[lo].a;
  compact_write_block(a, blockIdx.x - segs[lo].first);
}

// ---- exclusive scan of two uint64 arrays in place (block sums + fix-up) ----

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = (uint64_t)kScanThreads * kScanItems;

__device__ __forceinline__ void block_exclusive_scan2(uint64_t& x, uint64_t& y, uint64_t& tx,
                                                      uint64_t& ty) {
  __shared__ uint64_t sx[kScanThreads], sy[kScanThreads];
  const int t = threadIdx.x;
  sx[t] = x; sy[t] = y;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    uint64_t ax = 0, ay = 0;
    if (t >= off) { ax = sx[t - off]; ay = sy[t - off]; }
    __syncthreads();
    sx[t] += ax; sy[t] += ay;
    __syncthreads();
  }
  tx = sx[kScanThreads - 1]; ty = sy[kScanThreads - 1];
  const uint64_t ix = sx[t], iy = sy[t];
  __syncthreads();
  x = ix - x; y = iy - y;
}

static __global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint64_t* a, const uint64_t* b,
                                                              uint64_t n, uint64_t* sums) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint64_t x = 0, y = 0;
  for (int i = 0; i < kScanItems; ++i) {
    const uint64_t j = base + (uint64_t)threadIdx.x * kScanItems + i;
    if (j < n) { x += a[j]; y += b[j]; }
  }
  uint64_t tx, ty;
  block_exclusive_scan2(x, y, tx, ty);
  if (threadIdx.x == 0) { sums[2 * blockIdx.x] = tx; sums[2 * blockIdx.x + 1] = ty; }
}

// Single block: exclusive scan over block sums; totals to sums[2*blocks..].
static __global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint64_t* sums, uint64_t blocks) {
  __shared__ uint64_t carry[2];
  if (threadIdx.x == 0) { carry[0] = 0; carry[1] = 0; }
  __syncthreads();
  for (uint64_t start = 0; start < blocks; start += kScanThreads) {
    const uint64_t j = start + threadIdx.x;
    uint64_t x = 0, y = 0;
    if (j < blocks) { x = sums[2 * j]; y = sums[2 * j + 1]; }
    uint64_t tx, ty;
    block_exclusive_scan2(x, y, tx, ty);
    const uint64_t cx = carry[0], cy = carry[1];
    if (j < blocks) { sums[2 * j] = x + cx; sums[2 * j + 1] = y + cy; }
    __syncthreads();
    if (threadIdx.x == 0) { carry[0] = cx + tx; carry[1] = cy + ty; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { sums[2 * blocks] = carry[0]; sums[2 * blocks + 1] = carry[1]; }
}

static __global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint64_t* a, uint64_t* b, uint64_t n,
                                                             const uint64_t* sums) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint64_t va[kScanItems], vb[kScanItems];
  uint64_t x = 0, y = 0;
  for (int i = 0; i < kScanItems; ++i) {
    const uint64_t j = base + (uint64_t)threadIdx.x * kScanItems + i;
    va[i] = j < n ? a[j] : 0; vb[i] = j < n ? b[j] : 0;
    x += va[i]; y += vb[i];
  }
  uint64_t tx, ty;
  block_exclusive_scan2(x, y, tx, ty);
  x += sums[2 * blockIdx.x]; y += sums[2 * blockIdx.x + 1];
  for (int i = 0; i < kScanItems; ++i) {
    const uint64_t j = base + (uint64_t)threadIdx.x * kScanItems + i;
    if (j < n) { a[j] = x; b[j] = y; }
    x += va[i]; y += vb[i];
  }
}

}  // namespace dpah
