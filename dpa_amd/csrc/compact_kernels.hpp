// compact_kernels.hpp -- K_compact: prune_integral_curves
// (particle_advector.cpp:402-408), index_generator::generate
// (index_generator.cpp:10-39) and color_generator::generate_from_angular_velocities
// (color_generator.cpp:37-58) fused over one chunk of a round's strided slots.
//
// Slot row p (stride s) holds particle p's seed, one vertex per step and the
// terminal vertex at k_p + 1; rows are never pre-filled, so the kernels read
// only slots [0, k_p + 1].  For a row with no sentinel-valued vertex the pruned
// curve is the whole row, it contributes k_p + 2 vertices and k_p segments
// (b + j, b + j + 1), j < k_p.  A row holding a vertex equal to (-1,-1,-1) or
// (-2,-2,-2) -- dpa would drop / treat it as terminal -- is walked serially with
// the literal rules.
#pragma once

#include "device_math.hpp"

namespace dpah {

struct CompactArgs {
  const float*    slots;     // [count][stride][3]
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
  const float* row = a.slots + 3 * p * a.stride;
  uint64_t kept = 0, pairs = 0;
  bool have_prev = false, prev_term = false;
  for (uint64_t j = 0; j <= k + 1; ++j) {
    const float* v = row + 3 * j;
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
  const float* row = a.slots + 3 * p * a.stride;
  const uint64_t vb = a.vcount[p], ib = a.icount[p];
  // list of kept slot indices is walked twice with a small lookahead
  uint64_t m = 0, pairs = 0;
  int64_t prev_j = -1;
  for (uint64_t j = 0; j <= k + 1; ++j) {
    const float* v = row + 3 * j;
    if (eq3(v, -1.0f)) continue;
    // next kept slot
    int64_t next_j = -1;
    for (uint64_t q = j + 1; q <= k + 1; ++q)
      if (!eq3(row + 3 * q, -1.0f)) { next_j = (int64_t)q; break; }
    const uint64_t g = a.vertex_base + vb + m;
    float* out = a.vertices + 3 * (vb + m);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    float col = 0.0f;
    const bool term = eq3(v, -2.0f);
    if (!term) {
      float prev[3] = {0.0f, 0.0f, 0.0f}, next[3] = {0.0f, 0.0f, 0.0f};
      if (g > 1 && prev_j >= 0 && !eq3(row + 3 * prev_j, -2.0f)) {
        const float* u = row + 3 * prev_j;
        normalized3(v[0] - u[0], v[1] - u[1], v[2] - u[2], prev);
      }
      if (next_j >= 0 && !eq3(row + 3 * next_j, -2.0f)) {
        const float* u = row + 3 * next_j;
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

// Writes: one wave per row, lanes stride over the row's slots.
static __global__ __launch_bounds__(256) void k_compact_write(CompactArgs a) {
  const uint64_t p = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  if (p >= a.count) return;
  if (a.special[p]) {
    if (lane == 0) compact_special_row(a, p);
    return;
  }
  const uint64_t k = a.steps[p];
  const float* row = a.slots + 3 * p * a.stride;
  const uint64_t vb = a.vcount[p], ib = a.icount[p];
  for (uint64_t j = lane; j <= k + 1; j += 64) {
    const float* v = row + 3 * j;
    const float x = v[0], y = v[1], z = v[2];
    float* out = a.vertices + 3 * (vb + j);
    out[0] = x; out[1] = y; out[2] = z;
    float col = 0.0f;
    if (j <= k) {
      const uint64_t g = a.vertex_base + vb + j;
      float prev[3] = {0.0f, 0.0f, 0.0f}, next[3] = {0.0f, 0.0f, 0.0f};
      if (j >= 1 && g > 1) {
        const float* u = v - 3;
        normalized3(x - u[0], y - u[1], z - u[2], prev);
      }
      if (j + 1 <= k) {
        const float* u = v + 3;
        normalized3(u[0] - x, u[1] - y, u[2] - z, next);
        put_index(a, ib + j, g);
      }
      col = angle_color(prev, next);
    }
    a.colors[vb + j] = col;
  }
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
