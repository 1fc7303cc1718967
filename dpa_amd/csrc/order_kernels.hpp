// order_kernels.hpp -- spatial ordering of the persistent kernel's work queue.
//
// The queue hands particles out in the Morton order of the brick (a cube of
// 2^shift cells per axis) their start position lies in: a counting sort of
// brick keys (key + histogram, three-pass scan, scatter).  The persistent
// kernel splits the ordered queue into one contiguous part per XCD and a
// workgroup drains its own XCD's part before it takes from the others.  The
// 64 lanes of a wave then start in a few neighbouring bricks, and the particles
// one XCD's L2 serves cover one compact region, so a corner reload mostly hits
// lines a neighbouring lane fetched.  Particles are independent; the order
// changes which lane runs which particle when, never a result.
#pragma once

#include "advect_kernels.hpp"

namespace dpah {

constexpr uint32_t kOrderMaxBits = 6;      // brick coordinate bits per axis (2^18 keys)

__device__ __forceinline__ uint32_t spread_bits3(uint32_t v) {   // bit i -> bit 3i (10 bits)
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

struct OrderArgs {
  uint32_t shift;            // brick = 2^shift cells per axis
  uint32_t bits;             // brick coordinate bits per axis
  uint32_t keys;             // 2^(3 bits) + 1; the last key collects seeds outside the grid
  uint32_t* key;             // [n] key of queue ordinal q
  uint32_t* count;           // [keys] histogram, scanned to offsets in place
  uint32_t* order;           // [n] position in the ordered queue -> ordinal q
};

// Brick Morton key of ordinal q's start position.  The cell index is a plain
// float quotient: it only groups particles, it decides nothing about them.
__device__ __forceinline__ uint32_t brick_key(const AdvectArgs& a, const OrderArgs& o,
                                              uint64_t q) {
  float* row = nullptr;
  const uint64_t pid = resolve(a, q, row);
  float p0, p1, p2;
  seed_position(a, pid, p0, p1, p2);
  const float q0 = floorf((p0 - a.f.o0) / a.f.h0), q1 = floorf((p1 - a.f.o1) / a.f.h1),
              q2 = floorf((p2 - a.f.o2) / a.f.h2);
  const float lim = (float)(1u << (o.bits + o.shift));
  if (!(q0 >= 0.0f && q1 >= 0.0f && q2 >= 0.0f && q0 < lim && q1 < lim && q2 < lim))
    return o.keys - 1;
  const uint32_t b0 = (uint32_t)q0 >> o.shift, b1 = (uint32_t)q1 >> o.shift,
                 b2 = (uint32_t)q2 >> o.shift;
  return (spread_bits3(b0) << 2) | (spread_bits3(b1) << 1) | spread_bits3(b2);
}

static __global__ __launch_bounds__(256) void k_order_clear(OrderArgs o) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < o.keys; k += gridDim.x * blockDim.x)
    o.count[k] = 0u;
}

static __global__ __launch_bounds__(256) void k_order_key(AdvectArgs a, OrderArgs o) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const uint32_t k = brick_key(a, o, q);
  o.key[q] = k;
  atomicAdd(o.count + k, 1u);
}

// Exclusive scan of the histogram in three passes: tile totals, a one-block
// scan of the totals, then each tile scanned with its offset.
constexpr uint32_t kOrderScanItems = 8;
constexpr uint32_t kOrderTile = 256 * kOrderScanItems;

// Exclusive scan of one value per thread over a 256-thread block; returns the
// block total through `total`.
__device__ __forceinline__ uint32_t block_exclusive_scan_u32(uint32_t v, uint32_t& total) {
  __shared__ uint32_t sh[256];
  const uint32_t t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t x = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  total = sh[255];
  const uint32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

static __global__ __launch_bounds__(256) void k_order_tile_sums(OrderArgs o, uint32_t* sums) {
  const uint32_t base = blockIdx.x * kOrderTile + threadIdx.x * kOrderScanItems;
  uint32_t s = 0;
  for (uint32_t i = 0; i < kOrderScanItems; ++i)
    if (base + i < o.keys) s += o.count[base + i];
  uint32_t total;
  block_exclusive_scan_u32(s, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block over the tile totals (tiles <= 256 * kOrderScanItems).
static __global__ __launch_bounds__(256) void k_order_scan_sums(uint32_t* sums, uint32_t tiles) {
  const uint32_t base = threadIdx.x * kOrderScanItems;
  uint32_t v[kOrderScanItems], s = 0;
  for (uint32_t i = 0; i < kOrderScanItems; ++i) {
    v[i] = base + i < tiles ? sums[base + i] : 0u;
    s += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan_u32(s, total);
  for (uint32_t i = 0; i < kOrderScanItems; ++i) {
    if (base + i < tiles) sums[base + i] = run;
    run += v[i];
  }
}

static __global__ __launch_bounds__(256) void k_order_scan_apply(OrderArgs o, const uint32_t* sums) {
  const uint32_t base = blockIdx.x * kOrderTile + threadIdx.x * kOrderScanItems;
  uint32_t v[kOrderScanItems], s = 0;
  for (uint32_t i = 0; i < kOrderScanItems; ++i) {
    v[i] = base + i < o.keys ? o.count[base + i] : 0u;
    s += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan_u32(s, total) + sums[blockIdx.x];
  for (uint32_t i = 0; i < kOrderScanItems; ++i) {
    if (base + i < o.keys) o.count[base + i] = run;
    run += v[i];
  }
}

// Scatter: ordinals of one key land in an arbitrary order within the key's
// range, which is harmless (the order is a schedule, not a result).
static __global__ __launch_bounds__(256) void k_order_scatter(uint64_t n, OrderArgs o) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t slot = atomicAdd(o.count + o.key[q], 1u);
  if (slot < n) o.order[slot] = (uint32_t)q;   // always true for a correct scan
}

}  // namespace dpah
