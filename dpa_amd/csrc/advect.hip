// advect.hip -- launchers of the K_advect kernels (advect_kernels.hpp) and of
// the queue ordering that precedes a persistent launch (order_kernels.hpp).
#include <algorithm>
#include <type_traits>

#include "advect_kernels.hpp"
#include "order_kernels.hpp"
#include "runtime.hpp"

namespace dpah {

namespace {

template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv>
void launch_t(dpah_context* ctx, const Tuning& T, const AdvectArgs& a, bool persistent,
              hipStream_t s) {
  if (a.n == 0) return;
  if (persistent) {
    const int block = T.block;
    // recorded launches: 6 (each workgroup holds a 24 KB slot stage, and 6 fill
    // the 160 KB of LDS of a CU); endpoints: 4 for seed arrays; for lattices
    // of steady fields 3 on C-order nodes (their neighbours' node rows hit in
    // the L2), and on node bricks, whose lines hold more of what neighbouring
    // particles read, 5 for auxiliary lattices and 4 for lattices of points
    // (FTLE on ABC 512x512x128: +-1/2-cell 24.75 / 22.08 / 21.66 / 21.97 ms at
    // 3 / 4 / 5 / 6, dpa lattice 9.52 / 9.14 / 9.41 at 3 / 4 / 5); 4 for
    // lattices of time-variant fields, whose reloads read two slices (a
    // 256x256x16-point FTLE slab on a 256^3 x 16-slice field: 3.32 ms at 4,
    // 3.43 at 5, 3.55-3.70 at 3)
    const int bpc = T.blocks_per_cu ? T.blocks_per_cu
                  : kRecord ? 6
                  : a.seed_kind == kSeedArray || kUnsteady ? 4
                  : !a.f.nbrick ? 3
                  : a.seed_kind == kSeedLattice ? 4 : 5;
    uint64_t grid = (uint64_t)ctx->cu_count * bpc;
    const uint64_t need = (a.n + block - 1) / block;
    if (grid > need) grid = need;
    const bool lattice = a.seed_kind != kSeedArray;
    DPAH_REQUIRE(!(kRecord && lattice), "lattice-seeded persistent launches record no curves");
    ctx->timed("advect_persistent", s, [&] {
      // the load-counting instance exists for steady dpa-mode endpoint launches
      constexpr bool kCountable = !kRecord && kIntegrator == kDpaRk4 && !kUnsteady;
      if constexpr (!kRecord) {
        if (lattice) {
          // (an instance that assumes the node layout with 32-bit offsets
          // measured 58.39 -> 58.20 ms on the auxiliary FTLE launch and 10.87
          // -> 11.00 on the dpa-lattice one: not kept)
          if (kCountable && a.loads_out)
            hipLaunchKernelGGL((k_advect_persistent<false, kIntegrator, kUnsteady, kDiv,
                                                    kSlotDirect, kCountable, true>),
                               dim3((unsigned)grid), dim3(block), 0, s, a);
          else
            hipLaunchKernelGGL((k_advect_persistent<false, kIntegrator, kUnsteady, kDiv,
                                                    kSlotDirect, false, true>),
                               dim3((unsigned)grid), dim3(block), 0, s, a);
          return;
        }
      }
      // seed arrays on a cell layout read as one buffer resource (every
      // benchmark field) run instances that assume so
      auto go = [&](auto layout) {
        constexpr int kC = decltype(layout)::value;
        if (kRecord && a.slot_shift == kStagedShift)
          hipLaunchKernelGGL((k_advect_persistent<kRecord, kIntegrator, kUnsteady, kDiv,
                                                  kRecord ? kSlotDeferred : kSlotDirect, false,
                                                  false, kC>),
                             dim3((unsigned)grid), dim3(block), 0, s, a);
        else if (kCountable && a.loads_out)
          hipLaunchKernelGGL((k_advect_persistent<kRecord, kIntegrator, kUnsteady, kDiv,
                                                  kSlotDirect, kCountable, false, kC>),
                             dim3((unsigned)grid), dim3(block), 0, s, a);
        else
          hipLaunchKernelGGL((k_advect_persistent<kRecord, kIntegrator, kUnsteady, kDiv,
                                                  kSlotDirect, false, false, kC>),
                             dim3((unsigned)grid), dim3(block), 0, s, a);
      };
      if (a.f.cells && a.f.cell_bytes32) go(std::integral_constant<int, kLayoutCells32>{});
      else                        go(std::integral_constant<int, kLayoutAny>{});
    });
  } else {
    const uint64_t grid = (a.n + 255) / 256;
    ctx->timed("advect_simple", s, [&] {
      hipLaunchKernelGGL((k_advect_simple<kRecord, kIntegrator, kUnsteady, kDiv>), dim3((unsigned)grid),
                         dim3(256), 0, s, a);
    });
  }
  DPAH_HIP(hipGetLastError());
}

template <bool kRecord, int kDiv>
void dispatch_integrator(dpah_context* ctx, const Tuning& T, const AdvectArgs& a, int integrator,
                         bool unsteady, bool persistent, hipStream_t s) {
  if (integrator == kRk4) {
    if (unsteady) launch_t<kRecord, kRk4, true, kDiv>(ctx, T, a, persistent, s);
    else          launch_t<kRecord, kRk4, false, kDiv>(ctx, T, a, persistent, s);
  } else {
    if (unsteady) launch_t<kRecord, kDpaRk4, true, kDiv>(ctx, T, a, persistent, s);
    else          launch_t<kRecord, kDpaRk4, false, kDiv>(ctx, T, a, persistent, s);
  }
}

template <bool kRecord>
void dispatch_division(dpah_context* ctx, const Tuning& T, const AdvectArgs& a, int integrator,
                       bool unsteady, bool persistent, hipStream_t s) {
  switch (div_mode(a.f.exactdiv, a.f.fastdiv, unsteady)) {
    case kDivExact: dispatch_integrator<kRecord, kDivExact>(ctx, T, a, integrator, unsteady, persistent, s); break;
    case kDivFast:  dispatch_integrator<kRecord, kDivFast>(ctx, T, a, integrator, unsteady, persistent, s); break;
    default:        dispatch_integrator<kRecord, kDivide>(ctx, T, a, integrator, unsteady, persistent, s); break;
  }
}

// Clears the step counter and the queue heads (a kernel, not a memset, so the
// sequence stays valid inside a captured graph).
__global__ void k_reset_counters(unsigned long long* c) {
  const unsigned t = threadIdx.x;
  if (t == 1 || (t >= (unsigned)kQueueHeads && t < (unsigned)(kQueueHeads + kMaxParts)))
    c[t] = 0ull;
}

__global__ void k_clear_loads(unsigned long long* c) {
  if (threadIdx.x < 2) c[kLoadsCells + threadIdx.x] = 0ull;
}

// Device scratch of one ordering; released when the launch is enqueued, which
// is safe because the caching allocator hands blocks out again only in the
// context stream's order.
struct QueueOrder {
  DeviceBuffer key, count, order;
};

// Queue parts start on multiples of 64, so waves that draw 64 positions at a
// time draw whole aligned tiles of record slots.
void set_parts(AdvectArgs& a, uint32_t parts) {
  a.parts = parts;
  const uint64_t tiles = (a.n + kTileRows - 1) / kTileRows;
  for (uint32_t p = 0; p <= parts; ++p)
    a.part_lo[p] = std::min(a.n, (tiles / parts * p + std::min<uint64_t>(p, tiles % parts)) *
                                     kTileRows);
}

}  // namespace

// Brick size and key count of a field's queue ordering: bricks of 2^shift
// cells per axis (at least 2^min_shift; 0 asks for 4-cell bricks), at most
// 2^kOrderMaxBits bricks per axis; one more key for seeds outside the grid;
// tiles of the histogram scan.
void order_shape(const FieldView& f, uint32_t min_shift, uint32_t& shift, uint32_t& bits,
                 uint32_t& keys, uint32_t& tiles) {
  const int64_t D = std::max({(int64_t)f.n0, (int64_t)f.n1, (int64_t)f.n2});
  shift = min_shift ? min_shift : 2;
  while (((D + (1ll << shift) - 1) >> shift) > (1ll << kOrderMaxBits)) ++shift;
  bits = 1;
  while ((1ll << (bits + shift)) < D) ++bits;
  keys = (1u << (3 * bits)) + 1;
  tiles = (keys + kOrderTile - 1) / kOrderTile;   // <= 129 for 6 bits
}

// Seed arrays sorted by brick Morton key before the launch.  Endpoint
// launches only: a recorded launch writes one slot row per particle, and with
// the queue in seed order a wave's rows are neighbours in memory (2^18 seeds of
// the 256^3 vortex field x 1000 recorded steps: 4.0 ms in seed order, 5.1 ms ordered).
bool ordered_launch(const dpah_context* ctx, const Tuning& T, const AdvectLaunch& L) {
  return L.persistent && T.order && (ctx->caching || L.order_scratch) &&
         (L.stream == nullptr || L.order_scratch) && L.nseg == 0 &&
         L.slots == nullptr && L.seed_kind == kSeedArray && L.n >= kOrderMinParticles &&
         L.n < (1ull << 32);
}

// Lattices need no sort: the kernel unranks queue positions into 4x4x4 bricks
// of lattice points (lattice_brick_index), a point's auxiliary particles
// consecutive (six, or three when shared; the shared lattice's face particles
// follow in order).  A lattice in C order gives a wave 64 points along
// one row; bricks give it a 4x4x4 block (dpa-mode FTLE on 512x512x128: 16.1 ms
// advection, 39.7 ms in C order).  Whole lattices in one launch only, since
// the walk is over the whole lattice and the outputs are indexed by lattice
// point.
// nbrick: the field's nodes are read from node bricks (lattices of points walk
// smaller bricks in larger groups there: dpa lattice FTLE at 4 workgroups per
// CU 9.14 ms with 4-point bricks in groups of 2, 8.91 ms with 2-point bricks
// in groups of 4).
uint32_t lattice_brick(const Tuning& T, int seed_kind, bool nbrick) {
  if (T.lattice_brick) return (uint32_t)T.lattice_brick;
  return seed_kind == kSeedLattice ? (nbrick ? kLatticeBrickNb : kLatticeBrick)
         : seed_kind == kSeedAuxShared ? kAuxSharedBrick : kAuxBrick;
}

// Bricks per group edge of the grouped walk; 1: no groups.
uint32_t lattice_group(const Tuning& T, int seed_kind, bool nbrick) {
  if (T.lattice_group) return (uint32_t)T.lattice_group;
  return seed_kind == kSeedLattice ? (nbrick ? kLatticeGroupNb : kLatticeGroup) : 1u;
}

bool brick_launch(const Tuning& T, const AdvectLaunch& L) {
  if (!(L.persistent && T.order && L.nseg == 0 && L.slots == nullptr && L.lat.first == 0))
    return false;
  const uint64_t points = L.lat.d0 * L.lat.d1 * L.lat.d2;
  // the walk's 32-bit arithmetic: positions and a whole slab of bricks / groups
  const bool nb = L.f->nbrick != nullptr;
  if ((uint64_t)lattice_brick(T, L.seed_kind, nb) * lattice_group(T, L.seed_kind, nb) * L.lat.d1 *
          L.lat.d2 >= (1ull << 32))
    return false;
  if (L.seed_kind == kSeedLattice) return L.n == points && L.n < (1ull << 32);
  if (L.seed_kind == kSeedAux) return L.n == 6 * points && L.n < (1ull << 32);
  if (L.seed_kind == kSeedAuxShared)
    return L.n == aux_shared_count(L.lat.d0, L.lat.d1, L.lat.d2) && L.n < (1ull << 32);
  return false;
}

namespace {

void order_queue(dpah_context* ctx, const Tuning& T, AdvectArgs& a, QueueOrder& q,
                 const OrderScratch* keep, hipStream_t s) {
  OrderArgs o{};
  uint32_t tiles = 0;
  order_shape(a.f, (uint32_t)T.order_brick, o.shift, o.bits, o.keys, tiles);
  if (keep) {   // caller-owned buffers (a captured pass replays with the same ones)
    DPAH_REQUIRE(keep->n_cap >= a.n && keep->count_cap >= (uint64_t)o.keys + tiles,
                 "order scratch too small");
    o.key = keep->key;
    o.order = keep->order;
    o.count = keep->count;
  } else {
    q.key = DeviceBuffer(ctx, a.n * sizeof(uint32_t));
    q.order = DeviceBuffer(ctx, a.n * sizeof(uint32_t));
    q.count = DeviceBuffer(ctx, ((uint64_t)o.keys + tiles) * sizeof(uint32_t));
    o.key = q.key.as<uint32_t>();
    o.order = q.order.as<uint32_t>();
    o.count = q.count.as<uint32_t>();
  }
  const unsigned clear_grid = std::min<unsigned>((o.keys + 255) / 256, 1024u);
  const unsigned grid = (unsigned)((a.n + 255) / 256);
  ctx->timed("order_queue", s, [&] {
    hipLaunchKernelGGL(k_order_clear, dim3(clear_grid), dim3(256), 0, s, o);
    hipLaunchKernelGGL(k_order_key, dim3(grid), dim3(256), 0, s, a, o);
    uint32_t* sums = o.count + o.keys;
    hipLaunchKernelGGL(k_order_tile_sums, dim3(tiles), dim3(256), 0, s, o, sums);
    hipLaunchKernelGGL(k_order_scan_sums, dim3(1), dim3(256), 0, s, sums, tiles);
    hipLaunchKernelGGL(k_order_scan_apply, dim3(tiles), dim3(256), 0, s, o, sums);
    hipLaunchKernelGGL(k_order_scatter, dim3(grid), dim3(256), 0, s, a.n, o);
  });
  DPAH_HIP(hipGetLastError());
  a.order = o.order;
  set_parts(a, (uint32_t)std::max(1, std::min(T.parts, kMaxParts)));
}

}  // namespace

void clear_load_counters(dpah_context* ctx) {
  hipLaunchKernelGGL(k_clear_loads, dim3(1), dim3(64), 0, ctx->stream, ctx->d_counters);
  DPAH_HIP(hipGetLastError());
}

bool use_cells(const Tuning& T, bool scattered) {
  return T.cells == 2 || (T.cells == 1 && scattered);
}

std::string launch_key(const FieldView& f, int integrator, int seed_kind, bool record) {
  return std::to_string(f.n0) + "x" + std::to_string(f.n1) + "x" + std::to_string(f.n2) + "x" +
         std::to_string(f.nt) + ":i" + std::to_string(integrator) + ":s" +
         std::to_string(seed_kind) + ":r" + std::to_string(record ? 1 : 0);
}

const Tuning& tuning_for(const dpah_context* ctx, const std::string& key) {
  auto it = ctx->tuned.find(key);
  return it == ctx->tuned.end() ? ctx->tune : it->second;
}

uint32_t record_slot_shift(const dpah_context* ctx, const FieldView& f, int integrator,
                           uint64_t rows, bool persistent) {
  const Tuning& T = tuning_for(ctx, launch_key(f, integrator, kSeedArray, true));
  return slot_shift_for(rows, persistent, T.slot_stage);
}

bool record_grid(const dpah_context* ctx, const FieldView& f, int integrator, bool batch) {
  const Tuning& T = tuning_for(ctx, launch_key(f, integrator, kSeedArray, true));
  return recorded_grid(T.slot_stage, batch);
}

void launch_advect(dpah_context* ctx, const AdvectLaunch& L) {
  const Tuning& T =
      tuning_for(ctx, launch_key(*L.f, L.integrator, L.seed_kind, L.slots != nullptr));
  AdvectArgs a{};
  a.f = *L.f;
  if (!use_cells(T, L.seed_kind == kSeedArray)) a.f.cells = nullptr;
  if (!T.fastdiv) a.f.fastdiv = a.f.exactdiv = 0;
  a.segs = L.segs;
  a.nseg = L.nseg;
  a.seed_kind = L.seed_kind;
  a.lat = L.lat;
  a.seeds = L.seeds;
  a.rem_in = L.rem_in;
  a.uniform_rem = L.uniform_rem;
  a.pos_out = L.pos_out;
  a.rem_out = L.rem_out;
  a.status_out = L.status_out;
  a.steps_out = L.steps_out;
  a.special_out = L.special_out;
  a.slots = L.slots;
  a.stride = L.stride;
  a.n = L.n;
  a.t0 = L.t0;
  a.coef = make_coef(L.dt);
  unsigned long long* const counters = L.counters ? L.counters : ctx->d_counters;
  a.queue = counters + kQueueHeads;
  a.total_steps = counters + 1;
  a.order = nullptr;
  set_parts(a, 1);
  // auto: 8 for endpoint launches of seed arrays (keeps lanes busy); 64 for
  // lattices (keeps neighbouring particles together) and for recorded launches
  // (a wave's 64 lanes then write one aligned tile of slots per step)
  a.refill = (uint32_t)(T.refill ? T.refill
                                 : (L.seed_kind == kSeedArray && !L.slots ? 8 : 64));
  DPAH_REQUIRE(L.slot_shift == 0 || (L.slot_shift == kStagedShift && L.persistent),
               "staged slot blocks need a persistent launch");
  a.slot_shift = L.slots ? L.slot_shift : 0u;
  // always the context's counters (also for the allocation-free form, whose
  // own counters are the scratch array's): loads accumulate across launches
  a.loads_out = ctx->count_loads ? ctx->d_counters + (a.f.cells ? kLoadsCells : kLoadsNodes)
                                 : nullptr;
  hipStream_t const s = L.stream ? L.stream : ctx->stream;
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters);
  DPAH_HIP(hipGetLastError());
  QueueOrder q;
  if (ordered_launch(ctx, T, L)) {
    order_queue(ctx, T, a, q, L.order_scratch, s);
  } else if (brick_launch(T, L)) {
    a.lat.brick = lattice_brick(T, L.seed_kind, a.f.nbrick != nullptr);
    const uint32_t g = lattice_group(T, L.seed_kind, a.f.nbrick != nullptr);
    a.lat.group = g > 1 ? a.lat.brick * g : 0u;
    set_parts(a, (uint32_t)std::max(1, std::min(T.parts, kMaxParts)));
  }
  const bool unsteady = L.f->nt > 1;
  if (L.slots) dispatch_division<true>(ctx, T, a, L.integrator, unsteady, L.persistent, s);
  else         dispatch_division<false>(ctx, T, a, L.integrator, unsteady, L.persistent, s);
}

}  // namespace dpah
