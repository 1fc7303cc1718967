// serve_kernels.hpp -- the tick kernel of the serving engine.
//
// Continuous batching: every admitted request owns per-particle state in HBM
// (position, remaining, iteration, status, record slots).  One tick launch
// advances the particles of ALL admitted requests that share a field and an
// integrator by at most `budget` steps each; requests join and leave between
// ticks.  A particle
// advanced in several ticks goes through exactly the operations it would go
// through in one launch, so results do not depend on the mix.
#pragma once

#include "advect_kernels.hpp"

namespace dpah {

struct TickSeg {
  uint64_t qbase;           // first ordinal of this segment in the tick, a multiple of 64
  uint64_t count;           // particles
  float* pos;               // [count][3] state (seed before the first tick)
  uint64_t* rem;            // [count]
  uint32_t* it;             // [count] steps taken so far
  uint8_t* status;          // [count], kActive until retired
  uint8_t* special;         // [count] record: a vertex equals a sentinel
  float* slots;             // record: slot layout of slot_float, slot shift 0
  uint64_t stride;
  float dt, t0;
  unsigned int* active_left;          // particles of the request not yet retired
  unsigned long long* steps;          // particle-steps of the request
};

struct TickArgs {
  FieldView f;
  const TickSeg* segs;
  uint32_t nseg;
  uint64_t n;               // total ordinals (every segment padded to a multiple of 64)
  uint32_t budget;          // steps per particle per tick
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
// retired particles), summed per segment across the wave when the kernel ends:
// a few atomics per wave instead of two per particle on the same two addresses.
struct Tally {
  uint32_t seg;              // ~0u: nothing held
  uint32_t retired;
  uint64_t steps;
};

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

// One thread per ordinal of the launch: the particle runs until it retires or
// has taken `budget` steps, then its state goes back to HBM.  The hardware
// dispatches workgroups in ordinal order, so the 64 lanes of a wave hold 64
// consecutive particles of one request (spatially coherent for lattice seeds)
// until all of them are done.  Segments start at multiples of 64, so a wave
// never straddles two: the segment is wave-uniform and its pointers live in
// scalar registers, which keeps occupancy at the per-request kernel's level.
template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv>
static __global__ __launch_bounds__(256) void k_advect_tick(TickArgs a) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Tally tally{~0u, 0u, 0ull};
  if (q < a.n) {   // a.n is a multiple of 64: whole waves pass or fail together
    const uint32_t seg = __builtin_amdgcn_readfirstlane(find_seg(a, q));
    const TickSeg& s = a.segs[seg];
    const uint64_t i = q - s.qbase;
    if (i < s.count && s.status[i] == kActive) {
      Lane L;
      cache_clear(L.cc);
      L.pid = i;
      L.p0 = s.pos[3 * i]; L.p1 = s.pos[3 * i + 1]; L.p2 = s.pos[3 * i + 2];
      L.rem = s.rem[i];
      L.it = s.it[i];
      L.row = nullptr;
      L.special = false;
      if constexpr (kRecord) {
        L.row = s.slots + slot_float(i, s.stride, 0, 0);
        L.special = s.special[i] != 0;
      }
      const StepCoef coef = make_coef(s.dt);
      // budget >= 1 (checked on the host); the stop index wraps modulo 2^32,
      // and remaining < 2^32 - 2 retires the particle before a wrap matters
      const uint32_t stop = L.it + a.budget;
      uint8_t st;
      do {
        st = step_lane<kRecord, kIntegrator, kUnsteady, kDiv>(a.f, coef, s.t0, L);
      } while (st == kActive && L.it != stop);
      s.pos[3 * i] = L.p0; s.pos[3 * i + 1] = L.p1; s.pos[3 * i + 2] = L.p2;
      s.rem[i] = L.rem;
      tally.steps = L.it - (stop - a.budget);
      s.it[i] = L.it;
      tally.seg = seg;
      if (st != kActive) {
        if constexpr (kRecord) put3(L.row + slot_offset((uint64_t)L.it + 1, 0), -2.0f, -2.0f, -2.0f);
        s.status[i] = st;
        tally.retired = 1;
      }
      if constexpr (kRecord) s.special[i] = L.special ? 1 : 0;
    }
  }
  tally_flush_wave(a, tally);   // every lane of the wave takes part
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
    put3(a.slots + slot_float(i, a.stride, 0, 0), p0, p1, p2);
    a.special[i] = is_sentinel(p0, p1, p2) ? 1 : 0;
  }
}

}  // namespace dpah
