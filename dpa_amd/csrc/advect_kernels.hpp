// advect_kernels.hpp -- K_advect: the per-particle loop of
// particle_advector.cpp:325-396 as HIP kernels for gfx950.
//
// One lane integrates one particle at a time.  In the persistent form a lane
// whose particle retires (completed, left the grid, zero vector) takes the next
// particle from an on-device queue (one atomic per wave, ranks by mbcnt), so
// divergent retirement never idles a wave while work is left.  Particle state
// lives in registers; positions are AoS float3 in HBM, which 64 lanes read and
// write as one contiguous 768-byte run.
#pragma once

#include "device_math.hpp"

namespace dpah {

// Where a particle's start position comes from.
enum SeedKind : int {
  kSeedArray = 0,    // seeds[pid]
  kSeedLattice = 1,  // lattice point pid in C order (uniform_seed_generator.cpp:33-36)
  kSeedAux = 2,      // pid = 6*point + a: lattice point +- delta along axis a/2
  kSeedAuxShared = 3,  // delta = stride / 2: each particle shared by two points
                       // is advected once (aux_shared_particle)
};

// Field layouts a persistent kernel instance may assume (k_advect_persistent):
// kLayoutCells32, a cell layout read as one buffer resource (cell_bytes32).
enum : int { kLayoutAny = 0, kLayoutCells32 = 1 };

struct SeedLattice {
  float lo0, lo1, lo2;        // offset
  float st0, st1, st2;        // stride (physical)
  uint64_t d0, d1, d2;        // points per axis
  float dl0, dl1, dl2;        // auxiliary offsets (kSeedAux)
  uint64_t first;             // lattice index of particle / point 0 (chunked launches)
  // whole-lattice endpoint launches: the queue walks the lattice points in
  // bricks of brick^3 points (lattice_brick_index), the six particles of an
  // auxiliary point together; 0: in C order
  uint32_t brick;
  // and, when nonzero, walks groups of group^3 points (a multiple of brick)
  // one after another, bricks inside a group (lattice_walk_index)
  uint32_t group;
};

// Queue position q -> C-order index of the lattice point it starts, when the
// queue walks a d0 x d1 x d2 lattice (fewer than 2^32 points, and B*d1*d2 below
// 2^32) in bricks of B^3 points: bricks in C order, points in C order inside a
// brick, the bricks on the far faces cut to the lattice.  Every brick before
// the last along an axis is whole, so each level of the walk is one division:
// the slab of bricks along axis 0, the row of bricks along axis 1 inside it,
// the brick along axis 2 inside that, then the point inside the brick.  With
// B = 4 a wave that draws 64 aligned positions starts a 4x4x4 block of
// neighbours.
__host__ __device__ inline uint64_t lattice_brick_index(uint32_t d0, uint32_t d1, uint32_t d2,
                                                        uint32_t B, uint32_t q) {
  const uint32_t slab = B * d1 * d2;
  const uint32_t b0 = q / slab;
  uint32_t r = q - b0 * slab;
  const uint32_t h0 = d0 - B * b0 < B ? d0 - B * b0 : B;
  const uint32_t row = h0 * B * d2;
  const uint32_t b1 = r / row;
  r -= b1 * row;
  const uint32_t h1 = d1 - B * b1 < B ? d1 - B * b1 : B;
  const uint32_t brick = h0 * h1 * B;
  const uint32_t b2 = r / brick;
  r -= b2 * brick;
  const uint32_t h2 = d2 - B * b2 < B ? d2 - B * b2 : B;
  const uint32_t i0 = r / (h1 * h2);
  r -= i0 * h1 * h2;
  const uint32_t i1 = r / h2;
  const uint32_t i2 = r - i1 * h2;
  return ((uint64_t)(B * b0 + i0) * d1 + (B * b1 + i1)) * (uint64_t)d2 + (B * b2 + i2);
}

// The walk with groups of G^3 points (G a multiple of B) above the bricks:
// groups in C order, the groups on the far faces cut to the lattice, and
// inside each group (an h0 x h1 x h2 lattice of its own) the brick walk
// above.  A stretch of consecutive queue positions -- what a part of the
// queue has in flight at once -- then covers a compact block of the lattice
// rather than a slab of brick rows.  G = 0: the brick walk alone.  Needs G
// d1 d2 below 2^32.
__host__ __device__ inline uint64_t lattice_walk_index(uint32_t d0, uint32_t d1, uint32_t d2,
                                                       uint32_t B, uint32_t G, uint32_t q) {
  // G = 0: one group, the whole lattice (one brick walk in the code either way)
  uint32_t h0 = d0, h1 = d1, h2 = d2, r = q, o0 = 0, o1 = 0, o2 = 0;
  if (G != 0) {
    const uint32_t slab = G * d1 * d2;
    const uint32_t g0 = q / slab;
    r = q - g0 * slab;
    h0 = d0 - G * g0 < G ? d0 - G * g0 : G;
    const uint32_t row = h0 * G * d2;
    const uint32_t g1 = r / row;
    r -= g1 * row;
    h1 = d1 - G * g1 < G ? d1 - G * g1 : G;
    const uint32_t block = h0 * h1 * G;
    const uint32_t g2 = r / block;
    r -= g2 * block;
    h2 = d2 - G * g2 < G ? d2 - G * g2 : G;
    o0 = G * g0; o1 = G * g1; o2 = G * g2;
  }
  const uint64_t local = lattice_brick_index(h0, h1, h2, B, r);
  if (G == 0) return local;
  const uint32_t l = (uint32_t)local;
  const uint32_t i0 = l / (h1 * h2);
  const uint32_t rem = l - i0 * h1 * h2;
  const uint32_t i1 = rem / h2;
  const uint32_t i2 = rem - i1 * h2;
  return ((uint64_t)(o0 + i0) * d1 + (o1 + i1)) * (uint64_t)d2 + (o2 + i2);
}

// Record slots: rows tiled by 64, each row's slots grouped in blocks of 2^g
// (g, the slot shift, is a launch-wide choice: 0 or kStagedShift).  Block b of
// row r (slots b 2^g .. b 2^g + 2^g - 1) lives at float
// 3 2^g ((r / 64 * B + b) * 64 + r % 64), B = ceil(stride / 2^g) blocks per
// row, so block b of the 64 rows of a tile is one contiguous run and slot j of
// a row is slot_offset(j, g) floats past its slot 0.  g = 0: slot j of a
// tile's 64 rows is one 768-byte run, which a wave stepping those rows writes
// with one store per step.  g = 3: blocks of 96 bytes (three whole 32-byte
// sectors), which recorded launches large enough to stream their slots to
// HBM stage in LDS and write whole: a retired row then leaves whole sectors
// untouched rather than 12-byte holes in every run.  Slot buffers hold whole
// tiles (tiled_rows rows) of whole blocks.
constexpr uint64_t kTileRows = 64;
constexpr uint64_t kSlotPitch = 3 * kTileRows;                   // g = 0: slot to slot
constexpr uint32_t kStagedShift = 3;
constexpr uint64_t kSlotGroup = 1ull << kStagedShift;            // staged: slots per block
constexpr uint64_t kBlockFloats = 3 * kSlotGroup;                // staged: one row's block
constexpr uint64_t kTileBlockFloats = kBlockFloats * kTileRows;  // staged: block b of a tile

__host__ __device__ inline uint64_t tiled_rows(uint64_t rows) {
  return (rows + kTileRows - 1) / kTileRows * kTileRows;
}

__host__ __device__ inline uint64_t slot_blocks(uint64_t stride, uint32_t g) {
  return (stride + (1ull << g) - 1) >> g;
}

// Floats of the slot region of `rows` rows of `stride` slots.
__host__ __device__ inline uint64_t slot_region_floats(uint64_t rows, uint64_t stride, uint32_t g) {
  return tiled_rows(rows) * slot_blocks(stride, g) * (3ull << g);
}

// Slot j of a row, relative to its slot 0.
__host__ __device__ inline uint64_t slot_offset(uint64_t j, uint32_t g) {
  return (j >> g) * (kSlotPitch << g) + (j & ((1ull << g) - 1)) * 3;
}

__host__ __device__ inline uint64_t slot_float(uint64_t row, uint64_t stride, uint64_t j,
                                               uint32_t g) {
  return (row / kTileRows * slot_blocks(stride, g) * kTileRows + row % kTileRows) * (3ull << g) +
         slot_offset(j, g);
}

// Slot s (block s >> g >= b0) of row r of a tile, as a float offset from the
// start of the tile's block b0: the compaction reads a chunk of slots relative
// to its first block.
__host__ __device__ inline uint32_t tile_slot_float(uint32_t s, uint32_t b0, uint32_t r,
                                                    uint32_t g) {
  return (((s >> g) - b0) * (uint32_t)kTileRows + r) * (3u << g) + (s & ((1u << g) - 1u)) * 3u;
}

// A contiguous run of particles with one slot layout: one dpa round of one job.
// The queue hands out ordinals q; segment s owns [qbase, qbase + tiled_rows(count))
// (qbase a multiple of 64, so a wave never spans two segments) and maps ordinal
// q < qbase + count to particle first + (q - qbase), whose slot row is row
// q - qbase of the tiled region at slots + slot_base.
struct Segment {
  uint64_t qbase;
  uint64_t first;
  uint64_t count;
  uint64_t slot_base;   // in floats
  uint64_t stride;      // slots per row
};

struct AdvectArgs {
  FieldView f;
  const Segment*  segs;         // nseg > 0: batched form; 0: one implicit segment
  uint32_t        nseg;
  int             seed_kind;
  SeedLattice     lat;
  const float*    seeds;        // [n][3] (kSeedArray)
  const uint64_t* rem_in;       // [n], or null -> uniform_rem for every particle
  uint64_t        uniform_rem;
  float*          pos_out;      // [n][3]
  uint64_t*       rem_out;      // [n] or null
  uint8_t*        status_out;   // [n] or null
  uint32_t*       steps_out;    // [n] or null
  uint8_t*        special_out;  // [n] or null (record): a vertex equals a sentinel
  float*          slots;        // record: slot layout of slot_float, slot 0 = seed
  uint64_t        stride;
  uint64_t        n;
  float           t0;
  StepCoef        coef;
  unsigned long long* queue;        // [parts]: next position to hand out in each part
  unsigned long long* total_steps;  // one atomic add per wave at exit
  const uint32_t* order;            // position -> queue ordinal, or null (identity)
  uint32_t        parts;            // queue parts (one per XCD), 1..8
  uint64_t        part_lo[9];       // part p covers positions [part_lo[p], part_lo[p+1])
  uint32_t        refill;           // a wave draws new particles once this many lanes idle
  uint32_t        slot_shift;       // record: slot blocks of 2^slot_shift (kStagedShift:
                                    // staged through LDS, persistent launches only)
  unsigned long long* loads_out;    // load-counting instance: cell cache reloads, one
                                    // atomic add per wave at exit
};

__device__ __forceinline__ bool is_sentinel(float a, float b, float c) {
  return (a == -1.0f && b == -1.0f && c == -1.0f) || (a == -2.0f && b == -2.0f && c == -2.0f);
}

__device__ __forceinline__ void put3(float* p, float a, float b, float c) {
  p[0] = a; p[1] = b; p[2] = c;
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ void add_wave_steps(unsigned long long* total, uint64_t mine) {
  // wave64 butterfly sum, then one atomic from lane 0
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off, 64);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(total, (unsigned long long)mine);
}

// Particle registers carried across the step loop.
struct Lane {
  CellCache cc;    // corners of the last sampled cell
  float* row;      // record: this particle's slot row (its slot 0)
  float* stage;    // staged record: this lane's slot 0 in the workgroup's LDS stage
  uint64_t pid;
  uint64_t rem;
  uint32_t it;
  float p0, p1, p2;
  bool special;
};

// Slot writes.  kSlotDirect stores each slot as it comes (slot shift 0).  The staged modes
// (recorded persistent launches) keep a lane's current block in LDS --
// slot-major over the wave, slot s of lane l at float (s * 64 + l) * 3 of the
// wave's stage, so the 64 lanes' 12-byte accesses fall on distinct banks --
// and write the block's 96 bytes at once: kSlotStaged when the lane fills its
// block (or at the terminal vertex), kSlotDeferred leaves full blocks to the
// caller (flush_wave_block, one contiguous run for the whole wave).  Only a
// lane and its own wave touch its stage entries, in program order: no barrier.
// Slots past the terminal vertex get stale values; nothing reads a row past
// its slot k + 1, and the block lies inside the row.
enum SlotMode : int { kSlotDirect = 0, kSlotStaged = 1, kSlotDeferred = 2 };
constexpr uint32_t kStageFloatsPerWave = (uint32_t)(kSlotGroup * kTileRows * 3);

__device__ __forceinline__ void flush_block(const Lane& L, uint32_t b) {
  // slot by slot (three registers in flight): the eight 12-byte stores follow
  // one another, so each 32-byte sector is whole in L2 before it is written
  float* dst = L.row + (uint64_t)b * kTileBlockFloats;
#pragma unroll
  for (uint32_t q = 0; q < kSlotGroup; ++q) {
    const float* src = L.stage + q * (3 * kTileRows);
    put3(dst + 3 * q, src[0], src[1], src[2]);
  }
}

// Block b of the lanes in `mask`, for a wave whose lane l holds row 64t + l of
// one tile (tile = its rows' first block, block 0): the blocks of a tile are
// one 6144-byte run, written as 384 pieces of 16 bytes, piece m = lane + 64 q
// holding floats 4 (m % 6) .. +3 of row m / 6's block.  Every lane of the wave
// takes part; pieces of rows outside `mask` are skipped (whole sectors).
__device__ __forceinline__ void flush_wave_block(const float* wave_stage, float* tile, uint32_t b,
                                                 uint64_t mask, uint32_t lane) {
  float4* dst = (float4*)__builtin_assume_aligned(tile + (uint64_t)b * kTileBlockFloats, 16);
#pragma unroll 1
  for (uint32_t q = 0; q < kBlockFloats / 4; ++q) {   // one piece (4 registers) at a time
    const uint32_t m = lane + (uint32_t)kTileRows * q, r = m / 6, c = 4 * (m % 6);
    if ((mask >> r) & 1ull) {
      float v[4];
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t f = c + e;   // float f of the block: slot f / 3, component f % 3
        v[e] = wave_stage[((f / 3) * (uint32_t)kTileRows + r) * 3 + f % 3];
      }
      dst[m] = make_float4(v[0], v[1], v[2], v[3]);
    }
  }
}

template <int kSlots>
__device__ __forceinline__ void put_slot(const Lane& L, uint32_t j, float x, float y, float z) {
  if constexpr (kSlots == kSlotDirect) {
    put3(L.row + slot_offset(j, 0), x, y, z);   // direct writes: g = 0
  } else {
    put3(L.stage + (j % kSlotGroup) * (3 * kTileRows), x, y, z);
    if constexpr (kSlots == kSlotStaged)
      if (j % kSlotGroup == kSlotGroup - 1) flush_block(L, j / (uint32_t)kSlotGroup);
  }
}

// Multi-index of lattice point i (plus the launch's first point) in C order
// (indexing.hpp:10-20).
__device__ __forceinline__ void lattice_unravel(const SeedLattice& g, uint64_t i, uint64_t& i0,
                                                uint64_t& i1, uint64_t& i2) {
  i += g.first;
  i2 = i % g.d2;
  const uint64_t r = i / g.d2;
  i1 = r % g.d1;
  i0 = r / g.d1;
}

// position = offset + stride * multi_index, each product and sum rounded
// (uniform_seed_generator.cpp:35-36).
__device__ __forceinline__ void lattice_point(const SeedLattice& g, uint64_t i, float& p0,
                                              float& p1, float& p2) {
  uint64_t i0, i1, i2;
  lattice_unravel(g, i, i0, i1, i2);
  { const float m = g.st0 * (float)i0; p0 = g.lo0 + m; }
  { const float m = g.st1 * (float)i1; p1 = g.lo1 + m; }
  { const float m = g.st2 * (float)i2; p2 = g.lo2 + m; }
}

// The auxiliary particle of lattice point `point` moved by +delta (plus) or
// -delta along axis ax: that coordinate is lo + fl(st k +- delta), the offset
// position rounded once (an fmaf), the others the lattice point's.  With
// delta = st / 2 the +delta particle of point k and the -delta particle of
// point k + 1 are the same real number st k + delta, so the same float.
__device__ __forceinline__ void aux_particle(const SeedLattice& g, uint64_t point, uint32_t ax,
                                             bool plus, float& p0, float& p1, float& p2) {
  uint64_t i0, i1, i2;
  lattice_unravel(g, point, i0, i1, i2);
  { const float m = g.st0 * (float)i0; p0 = g.lo0 + m; }
  { const float m = g.st1 * (float)i1; p1 = g.lo1 + m; }
  { const float m = g.st2 * (float)i2; p2 = g.lo2 + m; }
  if (ax == 0) {
    const float m = __builtin_fmaf(g.st0, (float)i0, plus ? g.dl0 : -g.dl0); p0 = g.lo0 + m;
  } else if (ax == 1) {
    const float m = __builtin_fmaf(g.st1, (float)i1, plus ? g.dl1 : -g.dl1); p1 = g.lo1 + m;
  } else {
    const float m = __builtin_fmaf(g.st2, (float)i2, plus ? g.dl2 : -g.dl2); p2 = g.lo2 + m;
  }
}

// Particles of an auxiliary lattice whose offsets are half its stride, over
// d0 whole slabs along axis 0 (points first .. first + d0 d1 d2 - 1, first a
// multiple of d1 d2): particle 3 p + a (p < N = d0 d1 d2, chunk-local C order)
// is point p's +delta particle along axis a, which is also point p + e_a's
// -delta particle; then the -delta particles of the points on the chunk's
// low faces, which no point inside the chunk owns: axis 0 (d1 d2 of them, C
// order over (i1, i2)), axis 1 (d0 d2, over (i0, i2)), axis 2 (d0 d1, over
// (i0, i1)).
__host__ __device__ inline uint64_t aux_shared_count(uint64_t d0, uint64_t d1, uint64_t d2) {
  return 3 * d0 * d1 * d2 + d1 * d2 + d0 * d2 + d0 * d1;
}

__device__ __forceinline__ void aux_shared_particle(const SeedLattice& g, uint64_t pid, float& p0,
                                                    float& p1, float& p2) {
  const uint64_t n3 = 3 * g.d0 * g.d1 * g.d2;
  if (pid < n3) {
    const uint64_t point = pid / 3;
    aux_particle(g, point, (uint32_t)(pid - 3 * point), true, p0, p1, p2);
    return;
  }
  uint64_t f = pid - n3, point;
  uint32_t ax;
  if (f < g.d1 * g.d2) {                       // (0, i1, i2) = f
    ax = 0; point = f;
  } else if ((f -= g.d1 * g.d2) < g.d0 * g.d2) {   // (i0, 0, i2), f = i0 d2 + i2
    ax = 1; point = (f / g.d2) * g.d1 * g.d2 + f % g.d2;
  } else {                                     // (i0, i1, 0), f = i0 d1 + i1
    f -= g.d0 * g.d2;
    ax = 2; point = f * g.d2;
  }
  aux_particle(g, point, ax, false, p0, p1, p2);
}

// kLattice false: the launch seeds from an array (a.seed_kind is kSeedArray),
// and the lattice forms are compiled out of the kernel.
template <bool kLattice = true>
__device__ __forceinline__ void seed_position(const AdvectArgs& a, uint64_t pid, float& p0,
                                              float& p1, float& p2) {
  if (!kLattice || a.seed_kind == kSeedArray) {
    const float* s = a.seeds + 3 * pid;
    p0 = s[0]; p1 = s[1]; p2 = s[2];
  } else if (a.seed_kind == kSeedLattice) {
    lattice_point(a.lat, pid, p0, p1, p2);
  } else if (a.seed_kind == kSeedAux) {
    const uint64_t point = pid / 6;
    const uint32_t aux = (uint32_t)(pid - 6 * point);
    aux_particle(a.lat, point, aux >> 1, (aux & 1) != 0, p0, p1, p2);
  } else {
    aux_shared_particle(a.lat, pid, p0, p1, p2);
  }
}

// Queue ordinal -> particle index and slot row (binary search over segments;
// slot shift kShift, which is the launch's a.slot_shift); ~0 for the padding
// ordinals at the end of a segment's last tile.
template <uint32_t kShift = 0, bool kLattice = true>
__device__ __forceinline__ uint64_t resolve(const AdvectArgs& a, uint64_t q, float*& row) {
  if (a.nseg == 0) {
    row = a.slots + slot_float(q, a.stride, 0, kShift);
    if (kLattice && a.lat.brick) {   // brick-ordered lattice (whole-lattice endpoint launches)
      const uint32_t d0 = (uint32_t)a.lat.d0, d1 = (uint32_t)a.lat.d1, d2 = (uint32_t)a.lat.d2;
      if (a.seed_kind == kSeedAux) {
        const uint32_t point = (uint32_t)q / 6u;
        return 6 * lattice_walk_index(d0, d1, d2, a.lat.brick, a.lat.group, point) +
               ((uint32_t)q - 6u * point);
      }
      if (a.seed_kind == kSeedAuxShared) {   // points' particles walked, faces in order
        if ((uint32_t)q >= 3u * d0 * d1 * d2) return q;
        const uint32_t point = (uint32_t)q / 3u;
        return 3 * lattice_walk_index(d0, d1, d2, a.lat.brick, a.lat.group, point) +
               ((uint32_t)q - 3u * point);
      }
      return lattice_walk_index(d0, d1, d2, a.lat.brick, a.lat.group, (uint32_t)q);
    }
    return q;
  }
  uint32_t lo = 0, hi = a.nseg;            // last segment with qbase <= q
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.segs[mid].qbase <= q) lo = mid; else hi = mid;
  }
  const Segment& s = a.segs[lo];
  const uint64_t local = q - s.qbase;
  if (local >= s.count) return ~0ull;
  row = a.slots + s.slot_base + slot_float(local, s.stride, 0, kShift);
  return s.first + local;
}

template <bool kRecord, int kSlots = kSlotDirect, bool kLattice = true>
__device__ __forceinline__ void begin_particle(const AdvectArgs& a, Lane& L, uint64_t q) {
  float* row = nullptr;
  const uint64_t pid = resolve<kSlots == kSlotDirect ? 0u : kStagedShift, kLattice>(a, q, row);
  L.pid = pid;
  if (pid == ~0ull) return;   // segment padding: the lane stays idle
  cache_clear(L.cc);
  seed_position<kLattice>(a, pid, L.p0, L.p1, L.p2);
  L.rem = a.rem_in ? a.rem_in[pid] : a.uniform_rem;
  L.it = 0;
  if constexpr (kRecord) {
    L.row = row;
    L.special = is_sentinel(L.p0, L.p1, L.p2);
    put_slot<kSlots>(L, 0, L.p0, L.p1, L.p2);   // slot 0 never ends a block
  }
}

// kCountIt: the loop counted the particle's steps in L.it; otherwise they are
// its starting iterations minus the ones left.
// Staged: the caller has flushed every full block before the terminal vertex.
template <bool kRecord, bool kCountIt = true, int kSlots = kSlotDirect>
__device__ __forceinline__ uint64_t finish_particle(const AdvectArgs& a, const Lane& L, uint8_t st) {
  const uint64_t pid = L.pid;
  uint32_t it = L.it;
  if constexpr (!kCountIt) it = (uint32_t)((a.rem_in ? a.rem_in[pid] : a.uniform_rem) - L.rem);
  if constexpr (kRecord) {
    const uint32_t j = L.it + 1;   // the terminal vertex
    if constexpr (kSlots == kSlotDirect) {
      put_slot<kSlotDirect>(L, j, -2.0f, -2.0f, -2.0f);
    } else {
      put_slot<kSlotDeferred>(L, j, -2.0f, -2.0f, -2.0f);
      flush_block(L, j / (uint32_t)kSlotGroup);   // the row's last block, full or not
    }
    if (a.special_out) a.special_out[pid] = L.special ? 1 : 0;
  }
  put3(a.pos_out + 3 * pid, L.p0, L.p1, L.p2);
  if (a.rem_out) a.rem_out[pid] = L.rem;
  if (a.status_out) a.status_out[pid] = st;
  if (a.steps_out) a.steps_out[pid] = it;
  return it;
}

// One iteration of the loop at particle_advector.cpp:336-389 on field f with
// step coefficients c and start time t0.  Returns the retirement status, or
// kActive while the particle keeps going.
template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv, bool kCountIt = true,
          int kSlots = kSlotDirect>
__device__ __forceinline__ uint8_t step_lane(const FieldView& f, const StepCoef& c, float t0,
                                             Lane& L) {
  if (L.rem == 0) return kCompleted;
  float t = 0.0f;
  if constexpr (kUnsteady || kIntegrator == kRk4) {
    const float ft = (float)L.it * c.dt;   // iteration_index * step_size_
    t = t0 + ft;
  }
  V3 v;
  if (!sample<kUnsteady, kDiv>(f, L.p0, L.p1, L.p2, t, v, L.cc)) return kExited;
  if (is_zero(v)) return kZero;
  if constexpr (kIntegrator == kRk4) {
    if (!step_rk4<kUnsteady, kDiv>(f, L.p0, L.p1, L.p2, v, t, c, L.cc)) return kExited;
  } else {
    step_dpa(L.p0, L.p1, L.p2, v, c);
  }
  L.rem -= 1;
  if constexpr (kCountIt) L.it += 1;
  if constexpr (kRecord) {
    put_slot<kSlots>(L, L.it, L.p0, L.p1, L.p2);
    L.special |= is_sentinel(L.p0, L.p1, L.p2);
  }
  return L.rem == 0 ? kCompleted : kActive;
}

template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv>
__device__ __forceinline__ uint8_t step_once(const AdvectArgs& a, Lane& L) {
  return step_lane<kRecord, kIntegrator, kUnsteady, kDiv>(a.f, a.coef, a.t0, L);
}

// Per-request form: one thread per particle, runs it to retirement.
template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv>
__global__ __launch_bounds__(256) void k_advect_simple(AdvectArgs a) {
  const uint64_t pid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t mine = 0;
  if (pid < a.n) {
    Lane L;
    begin_particle<kRecord>(a, L, pid);
    if (L.pid != ~0ull) {
      uint8_t st;
      do {
        st = step_once<kRecord, kIntegrator, kUnsteady, kDiv>(a, L);
      } while (st == kActive);
      mine = finish_particle<kRecord>(a, L, st);
    }
  }
  add_wave_steps(a.total_steps, mine);
}

// Persistent form with the on-device work queue.  The queue positions
// [0, n) are split into `parts` contiguous parts, each with its own head;
// workgroups are dispatched to the XCDs round-robin, so workgroup b starts on
// part b % parts (its XCD's share of a spatially ordered queue) and moves on to
// the following parts once its own is drained.
// kSlots (recorded launches): kSlotDeferred stages slot blocks in LDS (slot
// shift kStagedShift), kSlotDirect stores each slot (slot shift 0).
// kCountLoads: also counts the register cell cache's reloads into
// a.loads_out (the roofline's byte count; a measurement instance, for steady
// dpa-mode endpoint launches: one sample per step, so a step reloaded exactly
// when it changed the cache's cell key).
// kLattice: the launch seeds a lattice (a.seed_kind kSeedLattice, kSeedAux or kSeedAuxShared,
// endpoint launches only); false compiles the lattice seeding and walk out of
// a seed-array launch's refill, where their uniform values would otherwise
// add to the scalar registers the step loop keeps.
// kLayout: what the launch checked of the field before it chose the
// instance -- kLayoutCells32, a cell layout read as one buffer resource;
// kLayoutAny, nothing.  The other reload paths drop out of the loop (256^3
// endpoints on cells: 369 -> 305 vector and 259 -> 217 scalar instructions of
// code, 100 -> 80 scalar registers).
template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv, int kSlots = kSlotDirect,
          bool kCountLoads = false, bool kLattice = false, int kLayout = kLayoutAny>
__global__ __launch_bounds__(256) void k_advect_persistent(AdvectArgs a) {
  static_assert(!kLattice || !kRecord, "lattice-seeded launches record no curves");
  // the field as the step loop sees it, with what the launch checked made
  // known to the compiler
  FieldView fv = a.f;
  if constexpr (kLayout == kLayoutCells32) {
    __builtin_assume(fv.cells != nullptr);
    __builtin_assume(fv.cell_bytes32 != 0);
  }
  // slot index (record) and time (time-variant fields, classic RK4) need the
  // step count inside the loop; steady dpa-mode endpoints only at the end
  constexpr bool kCountIt = kRecord || kUnsteady || kIntegrator == kRk4;
  static_assert(kRecord || kSlots == kSlotDirect, "endpoint launches write no slots");
  constexpr bool kStaged = kSlots != kSlotDirect;
  // staged launches: one stage of slot blocks per wave (4 waves of at most 256 threads)
  __shared__ float stage[kStaged ? 4 * kStageFloatsPerWave : 1];
  float* const wave_stage = stage + (threadIdx.x >> 6) * kStageFloatsPerWave;
  // Refill threshold 64 and no order table: a wave draws 64 aligned queue
  // positions at once, so lane l holds row 64t + l of one tile and the busy
  // lanes step in lockstep (same slot index): full blocks go out as one run
  // per wave.  Otherwise each lane writes its own block.
  const bool wave_runs = kStaged && a.refill == 64 && a.order == nullptr;
  const uint32_t lane = threadIdx.x & 63;
  Lane L;
  cache_clear(L.cc);
  L.row = nullptr;
  L.stage = wave_stage + 3 * lane;
  L.pid = ~0ull;
  L.rem = 0; L.it = 0; L.p0 = L.p1 = L.p2 = 0.0f; L.special = false;
  uint32_t part = blockIdx.x % a.parts;   // wave-uniform
  uint32_t parts_left = a.parts;          // parts this wave has not found drained
  bool open = true;                       // wave-uniform: some part may have particles
  static_assert(!kCountLoads || (!kRecord && !kUnsteady && kIntegrator == kDpaRk4),
                "load counting: steady dpa-mode endpoint launches");
  uint64_t mine = 0;
  uint64_t loads = 0;
  while (true) {
    const bool idle = (L.pid == ~0ull);
    if (open) {
      const uint64_t need = __ballot(idle);
      if (need && (uint32_t)__popcll(need) >= a.refill) {
        const uint64_t lo = a.part_lo[part], hi = a.part_lo[part + 1];
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.queue + part, (unsigned long long)cnt);
        base = __shfl(base, (int)leader, 64);
        if (idle) {
          const uint64_t pos = lo + base + lane_rank(need);
          const uint64_t q = pos < hi ? (a.order ? (uint64_t)a.order[pos] : pos) : a.n;
          if (q < a.n) begin_particle<kRecord, kSlots, kLattice>(a, L, q);   // q < n also guards the order table
        }
        if (lo + base + cnt >= hi) {
          part = part + 1 == a.parts ? 0 : part + 1;
          if (--parts_left == 0) open = false;
        }
      }
    }
    const bool busy = (L.pid != ~0ull);
    // a wave whose lanes all found a drained part keeps drawing from the next
    if (!__any(busy) && !open) break;
    if constexpr (!kStaged) {
      if (busy) {
        float k0 = 0.0f, k1 = 0.0f, k2 = 0.0f;
        if constexpr (kCountLoads) { k0 = L.cc.s0; k1 = L.cc.s1; k2 = L.cc.s2; }
        const uint8_t st =
            step_lane<kRecord, kIntegrator, kUnsteady, kDiv, kCountIt, kSlots>(fv, a.coef, a.t0, L);
        if constexpr (kCountLoads)
          loads += (L.cc.s0 != k0 || L.cc.s1 != k1 || L.cc.s2 != k2) ? 1u : 0u;
        if (st != kActive) {
          mine += finish_particle<kRecord, kCountIt, kSlots>(a, L, st);
          L.pid = ~0ull;
        }
      }
    } else {
      uint8_t st = kActive;
      bool full = false;   // this step's slot completed a block
      if (busy) {
        const uint32_t it0 = L.it;
        st = step_lane<kRecord, kIntegrator, kUnsteady, kDiv, kCountIt, kSlots>(fv, a.coef, a.t0, L);
        full = L.it != it0 && L.it % kSlotGroup == kSlotGroup - 1;
      }
      // full blocks go out before a terminal vertex reuses the stage
      if (wave_runs) {
        const uint64_t fm = __ballot(full);
        if (fm) {
          const int src = __ffsll((unsigned long long)fm) - 1;
          const uint32_t b = (uint32_t)__shfl((int)(L.it / (uint32_t)kSlotGroup), src, 64);
          const uint64_t t = (uint64_t)(L.row - kBlockFloats * lane);   // the tile's block 0
          const uint32_t tlo = (uint32_t)__shfl((int)(uint32_t)t, src, 64);
          const uint32_t thi = (uint32_t)__shfl((int)(uint32_t)(t >> 32), src, 64);
          float* tile = (float*)(((uint64_t)thi << 32) | (uint64_t)tlo);
          flush_wave_block(wave_stage, tile, b, fm, lane);
        }
      } else if (full) {
        flush_block(L, L.it / (uint32_t)kSlotGroup);
      }
      if (busy && st != kActive) {
        mine += finish_particle<kRecord, kCountIt, kSlots>(a, L, st);
        L.pid = ~0ull;
      }
    }
  }
  add_wave_steps(a.total_steps, mine);
  if constexpr (kCountLoads) add_wave_steps(a.loads_out, loads);
}

}  // namespace dpah
