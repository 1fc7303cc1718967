// serve_kernels.hpp -- the tick kernel of the serving engine.
//
// Continuous batching: every admitted request owns per-particle state in HBM
// (position, remaining, iteration, status, record slots).  One tick launch
// advances the particles of ALL admitted requests that share a field and an
// integrator by at most `budget` steps each, through the same on-device queue
// as the persistent kernel; requests join and leave between ticks.  A particle
// advanced in several ticks goes through exactly the operations it would go
// through in one launch, so results do not depend on the mix.
#pragma once

#include "advect_kernels.hpp"

namespace dpah {

struct TickSeg {
  uint64_t qbase;           // first queue ordinal of this segment in the tick
  uint64_t count;           // particles
  float* pos;               // [count][3] state (seed before the first tick)
  uint64_t* rem;            // [count]
  uint32_t* it;             // [count] steps taken so far
  uint8_t* status;          // [count], kActive until retired
  uint8_t* special;         // [count] record: a vertex equals a sentinel
  float* slots;             // record: [count][stride][3]
  uint64_t stride;
  float dt, t0;
  unsigned int* active_left;          // particles of the request not yet retired
  unsigned long long* steps;          // particle-steps of the request
};

struct TickArgs {
  FieldView f;
  const TickSeg* segs;
  uint32_t nseg;
  uint64_t n;               // total ordinals
  uint32_t budget;          // steps per particle per tick
  unsigned long long* queue;
};

__device__ __forceinline__ uint32_t find_seg(const TickArgs& a, uint64_t q) {
  uint32_t lo = 0, hi = a.nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.segs[mid].qbase <= q) lo = mid; else hi = mid;
  }
  return lo;
}

// Per-lane tallies of one segment's request counters (particle-steps and
// retired particles), flushed when the lane moves to another segment and,
// grouped by segment across the wave, when the kernel ends: a few atomics per
// wave instead of two per particle on the same two addresses.
struct Tally {
  uint32_t seg;              // ~0u: nothing held
  uint32_t retired;
  uint64_t steps;
};

__device__ __forceinline__ void tally_flush_lane(const TickArgs& a, Tally& t) {
  if (t.seg == ~0u) return;
  const TickSeg& s = a.segs[t.seg];
  if (t.steps) atomicAdd(s.steps, (unsigned long long)t.steps);
  if (t.retired) atomicSub(s.active_left, t.retired);
  t.seg = ~0u; t.steps = 0; t.retired = 0;
}

__device__ __forceinline__ void tally_flush_wave(const TickArgs& a, Tally& t) {
  const uint32_t lane = threadIdx.x & 63;
  while (true) {
    const uint64_t held = __ballot(t.seg != ~0u);
    if (!held) break;
    const int leader = __ffsll((unsigned long long)held) - 1;
    const uint32_t seg = __shfl(t.seg, leader, 64);
    const bool mine = t.seg == seg;
    uint64_t st = mine ? t.steps : 0;
    uint32_t rt = mine ? t.retired : 0;
    for (int off = 32; off > 0; off >>= 1) {
      st += __shfl_xor(st, off, 64);
      rt += __shfl_xor(rt, off, 64);
    }
    if ((int)lane == leader) {
      const TickSeg& s = a.segs[seg];
      if (st) atomicAdd(s.steps, (unsigned long long)st);
      if (rt) atomicSub(s.active_left, rt);
    }
    if (mine) { t.seg = ~0u; t.steps = 0; t.retired = 0; }
  }
}

template <bool kRecord, int kIntegrator, bool kUnsteady>
static __global__ __launch_bounds__(256) void k_advect_tick(TickArgs a) {
  const uint32_t lane = threadIdx.x & 63;
  Lane L;
  cache_clear(L.cc);
  L.row = nullptr;
  L.pid = ~0ull;
  L.rem = 0; L.it = 0; L.p0 = L.p1 = L.p2 = 0.0f; L.special = false;
  uint32_t seg = 0, left = 0;
  StepCoef coef = make_coef(0.0f);
  float t0 = 0.0f;
  uint32_t it_begin = 0;
  Tally tally{~0u, 0u, 0ull};
  bool open = true;
  while (true) {
    const bool idle = (L.pid == ~0ull);
    if (open) {
      const uint64_t need = __ballot(idle);
      if (need) {
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.queue, (unsigned long long)cnt);
        base = __shfl(base, (int)leader, 64);
        if (idle) {
          const uint64_t q = base + lane_rank(need);
          if (q < a.n) {
            seg = find_seg(a, q);
            const TickSeg& s = a.segs[seg];
            const uint64_t i = q - s.qbase;
            if (s.status[i] == kActive) {
              L.pid = i;
              cache_clear(L.cc);
              L.p0 = s.pos[3 * i]; L.p1 = s.pos[3 * i + 1]; L.p2 = s.pos[3 * i + 2];
              L.rem = s.rem[i];
              L.it = s.it[i];
              if constexpr (kRecord) {
                L.row = s.slots + 3 * i * s.stride;
                L.special = s.special[i] != 0;
              }
              coef = make_coef(s.dt);
              t0 = s.t0;
              left = a.budget;
              it_begin = L.it;
            }
          }
        }
        if (base + cnt >= a.n) open = false;
      }
    }
    const bool busy = (L.pid != ~0ull);
    // lanes that drew only retired particles draw again while the queue is open
    if (!__any(busy) && !open) {
      tally_flush_wave(a, tally);
      break;
    }
    if (busy) {
      const uint8_t st = step_lane<kRecord, kIntegrator, kUnsteady>(a.f, coef, t0, L);
      if (st == kActive) --left;              // budget >= 1 (checked on the host)
      if (st != kActive || left == 0) {
        const TickSeg& s = a.segs[seg];
        const uint64_t i = L.pid;
        s.pos[3 * i] = L.p0; s.pos[3 * i + 1] = L.p1; s.pos[3 * i + 2] = L.p2;
        s.rem[i] = L.rem;
        s.it[i] = L.it;
        if (tally.seg != seg) {
          tally_flush_lane(a, tally);
          tally.seg = seg;
        }
        tally.steps += L.it - it_begin;
        if (st != kActive) {
          if constexpr (kRecord) {
            put3(L.row + 3 * ((uint64_t)L.it + 1), -2.0f, -2.0f, -2.0f);
            s.special[i] = L.special ? 1 : 0;
          }
          s.status[i] = st;
          tally.retired += 1;
        } else if constexpr (kRecord) {
          s.special[i] = L.special ? 1 : 0;
        }
        L.pid = ~0ull;
      }
    }
  }
}

// State of a new request: positions from seeds (or auxiliary lattice points),
// remaining, it = 0, status active, slot 0 = seed for recorded requests.
struct TickInitArgs {
  const float* seeds;        // [count][3] or null (lattice in `lat`, kind kSeedAux)
  SeedLattice lat;
  int seed_kind;
  const uint64_t* rem_in;    // or null -> uniform
  uint64_t uniform_rem;
  uint64_t count;
  float* pos; uint64_t* rem; uint32_t* it; uint8_t* status; uint8_t* special;
  float* slots; uint64_t stride;   // record (segment-local rows)
};

static __global__ __launch_bounds__(256) void k_tick_init(TickInitArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.count) return;
  float p0, p1, p2;
  if (a.seed_kind == kSeedArray) {
    p0 = a.seeds[3 * i]; p1 = a.seeds[3 * i + 1]; p2 = a.seeds[3 * i + 2];
  } else {
    AdvectArgs tmp{};
    tmp.seed_kind = a.seed_kind;
    tmp.lat = a.lat;
    seed_position(tmp, i, p0, p1, p2);
  }
  a.pos[3 * i] = p0; a.pos[3 * i + 1] = p1; a.pos[3 * i + 2] = p2;
  a.rem[i] = a.rem_in ? a.rem_in[i] : a.uniform_rem;
  a.it[i] = 0;
  a.status[i] = kActive;
  if (a.slots) {
    put3(a.slots + 3 * i * a.stride, p0, p1, p2);
    a.special[i] = is_sentinel(p0, p1, p2) ? 1 : 0;
  }
}

}  // namespace dpah
