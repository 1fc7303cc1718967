// runtime.hpp -- internal state behind the C ABI: contexts, resident fields,
// results, error reporting and timed launches.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <new>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dpa_hip.h"
#include "advect_kernels.hpp"

namespace dpah {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DPAH_HIP(expr)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw ::dpah::Error(DPAH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

#define DPAH_REQUIRE(cond, msg)                                     \
  do {                                                              \
    if (!(cond)) throw ::dpah::Error(DPAH_ERR_INVALID, (msg));      \
  } while (0)

void set_error(const std::string& msg);
std::string& error_slot();   // the calling thread's last error message

#define DPAH_EXPORT extern "C" __attribute__((visibility("default")))

// Runs an entry point's body, turning exceptions into status codes.
template <class F>
int guarded(F&& body) {
  try {
    error_slot().clear();
    body();
    return DPAH_OK;
  } catch (const Error& e) {
    error_slot() = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    error_slot() = "host allocation failed";
    return DPAH_ERR_NOMEM;
  } catch (const std::exception& e) {
    error_slot() = e.what();
    return DPAH_ERR_INVALID;
  }
}

struct TimedKernel {
  uint64_t launches = 0;
  double total_ms = 0.0;
};

struct PendingTiming {
  std::string name;
  hipEvent_t start, stop;
};

// Launch-shape parameters picked by the autotuner (or defaults).
struct Tuning {
  int block = 256;          // threads per block of the persistent advect kernel
  // persistent grid = CUs * blocks_per_cu; 0: by launch kind (4 for endpoint
  // launches of seed arrays, 3 for lattices, 8 for recorded launches).  Fewer
  // particles in flight shrink what competes for each XCD's L2 and give each
  // lane more particles, against less latency hiding: 256^3 vortex endpoints
  // 5.04 ms at 8, 4.56 at 4, 4.95 at 3, 6.39 at 2; dpa-mode FTLE lattice
  // 14.8 ms at 8, 11.9 at 3; 64 recorded jobs 30.6 ms at 8, 30.2 at 6, 33.0 at 4
  int blocks_per_cu = 0;
  int order = 1;            // 1: spatially ordered queue split across the XCDs
  int parts = 8;            // queue parts (XCDs on the device)
  // idle lanes that make a wave draw new particles (1..64); 0: 8 for endpoint
  // launches of seed arrays (keeps lanes busy), 64 for lattices (keeps a wave's
  // neighbouring particles together) and recorded launches (aligned slot tiles)
  int refill = 0;
  // 0: node layout only; 1: fields also get a cell layout when it fits, read by
  // launches of seed arrays (scattered particles), while lattice seeds, whose
  // neighbouring lanes share node rows, read the nodes; 2: cells for every launch
  int cells = 1;
  int fastdiv = 1;          // reciprocal division on axes where it was verified exact
  // cell layout record order of fields created afterwards: 0 = C order, k =
  // bricks of 2^k cells per axis stored contiguously (8^3-cell bricks: the
  // 256^3 vortex workload 5.06 -> 5.01 ms, about 1%, the same for 4^3 and 16^3)
  int cell_brick = 3;
  // queue ordering of seed arrays: bricks of at least 2^order_brick cells per
  // axis (0: 4-cell bricks; larger when the grid needs more than 64 per axis)
  int order_brick = 0;
  // whole-lattice launches walk their points in bricks of this many points per
  // axis (2, 4 or 8); 0: by launch kind (kLatticeBrick for lattices of points,
  // kLatticeBrickNb on fields read from node bricks; kAuxBrick for auxiliary
  // lattices of six particles per point: FTLE ±½-cell 59.4 ms at 4, 58.4 at
  // 2; kAuxSharedBrick for shared particles: 28.6 ms at 4, 29.6 at 2, 29.8 at
  // 8 on C-order nodes, 21.85 / 22.35 / 22.74 on node bricks)
  int lattice_brick = 0;
  // and in groups of this many bricks per axis above them (2, 4, 8 or 16), so
  // that what a queue part has in flight is a compact block; 1: no groups;
  // 0: by launch kind (kLatticeGroup for lattices of points -- dpa-lattice
  // FTLE 11.47 ms without, 10.87 with groups of 2 -- and kLatticeGroupNb on
  // node bricks; none for auxiliary lattices, which lose 2-5% with groups)
  int lattice_group = 0;
  // recorded launches: 0 by launch kind (recorded_grid, slot_shift_for),
  // 1 persistent with one-slot runs, 2 persistent with 8-slot blocks staged in
  // LDS, 3 one thread per row (a grid over the rows) with one-slot runs
  int slot_stage = 0;
  // 1: steady fields created afterwards also keep their nodes in 4^3-node
  // bricks (the node bytes again) when they fit, read by launches that sample
  // nodes (FTLE +-1/2-cell lattice 28.8 -> 24.6 ms, dpa lattice 10.4 -> 9.4
  // ms); 2: time-variant fields too (the served 16-slice fields: no change at
  // saturation, 1.10-1.11e11 steps/s either way, for 12.9 GB more); 0: C-order
  // nodes only
  int node_brick = 1;
};

// Device counters of a context: [1] particle-steps, [2] FTLE misses,
// [3] division-check mismatches, [kLoadsCells] / [kLoadsNodes] cell cache
// reloads of load-counting launches that read cell records / nodes,
// [kQueueHeads, kQueueHeads + kMaxParts) the persistent queue's part heads.
constexpr int kCounterCount = 16;
constexpr int kLoadsCells = 4;
constexpr int kLoadsNodes = 5;
constexpr int kQueueHeads = 8;
constexpr int kMaxParts = 8;             // AdvectArgs::part_lo holds kMaxParts + 1

}  // namespace dpah

struct dpah_result;

struct dpah_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int xcds = 1;             // XCDs (accelerator complexes) on the device
  uint64_t memory_budget = 0;
  uint64_t slot_budget = 0;
  std::atomic<uint64_t> bytes_in_use{0};     // live blocks (rounded sizes)
  std::map<void*, uint64_t> user_allocs;      // dpah_device_alloc blocks and their sizes
  // Caching allocator: released blocks are kept, keyed by size, and handed out
  // again (stream order on the single context stream makes reuse safe), so
  // repeated passes do not pay hipMalloc / hipFree.  `caching` false makes
  // alloc / release go to the runtime every time (the per-request form).
  bool caching = true;
  std::multimap<uint64_t, void*> cached;      // size -> block
  std::map<void*, uint64_t> block_size;       // every live or cached block
  std::set<void*> direct;                     // live blocks taken while caching was off
  uint64_t cached_bytes = 0;
  void trim_cache();
  unsigned long long* d_counters = nullptr;   // kCounterCount counters (see Tuning)
  bool count_loads = false;   // endpoint launches count their cell cache reloads
  bool timing = false;
  std::map<std::string, dpah::TimedKernel> timings;
  std::vector<dpah::PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  dpah::Tuning tune;
  std::map<std::string, dpah::Tuning> tuned;   // autotuned launch key -> tuning
  std::string arch;              // gcnArchName
  std::string autotune_path;     // tuning cache file, empty: none
  std::map<uint64_t, bool> division_checked;   // (spacing bits, range bits) -> exact
  std::recursive_mutex mu;

  // Results freed by a thread that found the context busy (a long call holding
  // mu): deleted, and their blocks returned to the cache, at the next alloc
  // or drain_deferred, so a free never waits for another thread's GPU work.
  std::mutex deferred_mu;
  std::vector<dpah_result*> deferred;
  void drain_deferred();   // call with mu held
  // Pinned host memory for reading final states back: hipMemcpyAsync into
  // pageable memory goes through the runtime's small staging buffers at a few
  // GB/s and blocks the host meanwhile.  Grown on demand up to
  // kPinnedStagingMax (larger reads go to their pageable destination
  // directly); used with mu held and copied out before the call returns.
  static constexpr uint64_t kPinnedStagingMax = 256ull << 20;
  void* pinned = nullptr;
  uint64_t pinned_bytes = 0;
  void* host_staging(uint64_t bytes);   // nullptr above kPinnedStagingMax
  // Device-to-host copy on stream s that returns once the bytes are in dst
  // (call with mu held).  hipMemcpyAsync into pageable memory ran at 17 GB/s
  // (7.9 ms for 128 MiB into pages already touched); a pinned destination
  // takes one DMA (2.4 ms).  Pageable destinations of kStagedMin bytes and
  // more go through the pinned staging in kStagedChunk chunks, two in flight,
  // each chunk copied out by kStagedThreads host threads while the next one
  // transfers (2.7 ms).
  static constexpr uint64_t kStagedMin = 4ull << 20;
  static constexpr uint64_t kStagedChunk = 16ull << 20;
  static constexpr int kStagedThreads = 4;
  void download(void* dst, const void* src, uint64_t bytes, hipStream_t s);
  void* alloc(uint64_t bytes);
  void release(void* p, uint64_t bytes);
  hipEvent_t take_event();
  void resolve_timings();  // call after the stream is synchronised
  template <class F>
  void timed(const char* name, hipStream_t s, F&& launch) {
    if (!timing) { launch(); return; }
    dpah::PendingTiming t{name, take_event(), take_event()};
    DPAH_HIP(hipEventRecord(t.start, s));
    launch();
    DPAH_HIP(hipEventRecord(t.stop, s));
    pending.push_back(t);
  }
  template <class F>
  void timed(const char* name, F&& launch) { timed(name, stream, launch); }
  void sync() {
    DPAH_HIP(hipStreamSynchronize(stream));
    resolve_timings();
  }
};

struct dpah_field {
  dpah_context* ctx = nullptr;
  float* d_data = nullptr;
  uint64_t bytes = 0;
  float4* d_cells = nullptr;      // cell layout (FieldView::cells), or null
  uint64_t cell_bytes = 0;
  float* d_nbrick = nullptr;      // node bricks (FieldView::nbrick), or null
  uint64_t nbrick_bytes = 0;
  dpah::FieldView view{};
  uint64_t shape[3] = {0, 0, 0};
  uint64_t time_steps = 1;
};

namespace dpah {

struct Chunk {
  float* vertices = nullptr;   // device arrays (one block of `bytes`), or null once on the host
  float* colors = nullptr;
  void* indices = nullptr;
  uint64_t V = 0, I = 0;   // vertices, index entries
  uint64_t bytes = 0;
  // after materialize_host: the device block's bytes, uninitialised host memory
  std::shared_ptr<uint8_t> host;
  uint64_t colors_off = 0, indices_off = 0;   // offsets within the block
};

struct Round {
  std::vector<Chunk> chunks;
  uint64_t V = 0, I = 0;
};

}  // namespace dpah

struct dpah_result {
  dpah_context* ctx = nullptr;
  std::vector<dpah::Round> rounds;
  int use64 = 0;
  std::vector<float> final_pos;
  std::vector<uint64_t> final_rem;
  std::vector<uint8_t> final_status;
  uint64_t total_steps = 0;
  float* d_ftle = nullptr;        // FTLE requests: device volume
  uint64_t ftle_count = 0;
  std::shared_ptr<uint8_t> host_ftle;   // FTLE volume after materialize_host
  ~dpah_result();
};

namespace dpah {

// RAII device buffer owned by a context.
struct DeviceBuffer {
  dpah_context* ctx = nullptr;
  void* ptr = nullptr;
  uint64_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(dpah_context* c, uint64_t b) : ctx(c), ptr(c->alloc(b)), bytes(b) {}
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ctx(o.ctx), ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    reset();
    ctx = o.ctx; ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr;
    return *this;
  }
  ~DeviceBuffer() { reset(); }
  void reset() {
    if (ptr && ctx) ctx->release(ptr, bytes);
    ptr = nullptr;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Caller-owned scratch of the queue ordering (key, order: n_cap entries;
// count: count_cap entries).
struct OrderScratch {
  uint32_t* key = nullptr;
  uint32_t* order = nullptr;
  uint32_t* count = nullptr;
  uint64_t n_cap = 0, count_cap = 0;
};

// Particles worth ordering: enough to fill the persistent grid several times,
// and few enough for 32-bit positions.
constexpr uint64_t kOrderMinParticles = 1ull << 15;
constexpr uint32_t kLatticeBrick = 4;   // edge of a brick-ordered lattice queue, in points
constexpr uint32_t kAuxBrick = 2;       // the same for auxiliary lattices of six particles
constexpr uint32_t kAuxSharedBrick = 4; // and of shared particles (three per point)
constexpr uint32_t kLatticeGroup = 2;   // bricks per group edge, lattices of points
constexpr uint32_t kLatticeBrickNb = 2; // the same two on fields read from node bricks
constexpr uint32_t kLatticeGroupNb = 4;

// Launchers (advect.hip / compact.hip).
struct AdvectLaunch {
  const FieldView* f;
  int seed_kind;          // SeedKind
  SeedLattice lat;
  const Segment* segs;    // device table (batched form) or null
  uint32_t nseg;
  const float* seeds;
  const uint64_t* rem_in;
  uint64_t uniform_rem;
  float* pos_out;
  uint64_t* rem_out;
  uint8_t* status_out;
  uint32_t* steps_out;
  uint8_t* special_out;
  float* slots;
  uint64_t stride;
  uint32_t slot_shift = 0;   // slot_shift_for(rows the slot buffer was sized for)
  uint64_t n;
  float dt;
  float t0;
  int integrator;
  bool persistent;
  const OrderScratch* order_scratch = nullptr;   // null: temporaries from the allocator
  // step counter and queue heads (kCounterCount words); null: the context's
  unsigned long long* counters = nullptr;
  // stream of the launch; null: the context's.  On another stream a launch
  // sorts its queue only into caller-owned scratch (order_scratch): the
  // caching allocator reuses blocks in the context stream's order only.
  hipStream_t stream = nullptr;
};
void launch_advect(dpah_context* ctx, const AdvectLaunch& L);
// Brick shape of the queue ordering of a field (order_kernels.hpp).
void order_shape(const FieldView& f, uint32_t min_shift, uint32_t& shift, uint32_t& bits,
                 uint32_t& keys, uint32_t& tiles);
// Whether a launch sorts its queue (and so needs order scratch).
bool ordered_launch(const dpah_context* ctx, const Tuning& T, const AdvectLaunch& L);
// Whether a lattice launch walks its queue in closed-form bricks (no scratch).
bool brick_launch(const Tuning& T, const AdvectLaunch& L);
// The tuning a launch of this key uses (an autotuned entry or the context's).
const Tuning& tuning_for(const dpah_context* ctx, const std::string& key);
// Zeroes the cell-load counters (a kernel on the context stream).
void clear_load_counters(dpah_context* ctx);
// Whether a launch reads the cell layout (when its field has one).
bool use_cells(const Tuning& T, bool scattered);
// Autotuning key of a launch: field shape, integrator, seed kind, record flag.
std::string launch_key(const FieldView& f, int integrator, int seed_kind, bool record);
// Loads the context's autotune cache file (entries of this device) into ctx->tuned.
void load_autotune_cache(dpah_context* ctx);
void tuning_to_abi(const Tuning& t, dpah_tuning* out);
// Times launch() under every candidate tuning of `key` (prepare(t) runs untimed
// before each candidate's two timed launches), keeps the fastest in ctx->tuned
// and the autotune cache file, and returns it.
Tuning tune_key(dpah_context* ctx, const std::string& key, const std::vector<Tuning>& cands,
                const std::function<void(const Tuning&)>& prepare,
                const std::function<void()>& launch, double* best_ms);
// Looks up, once per process and device, the address of the allocation-free
// entry point's scratch array on `device` (which must be current).
void noalloc_setup(int device);

// Copies a result's device arrays to host memory and releases them (the
// per-request serving form returns results this way).
void materialize_host(dpah_result* r, hipStream_t stream);

// One rank's advection stage (dpah_advect); persistent=false uses one thread
// per particle (the per-request form).
dpah_result* run_advect(dpah_context* ctx, const dpah_field* field, const float* seeds,
                        const uint64_t* remaining, uint64_t uniform_remaining, uint64_t count,
                        const dpah_advect_params* params, bool persistent);

// The particles of an auxiliary-neighbour FTLE lattice of `shape` points:
// with delta half the lattice spacing on every axis (and every axis below
// 2^24 points) a particle between two neighbouring points is advected once
// (kSeedAuxShared: 3 per point plus the low faces' -delta particles), else six
// per point (kSeedAux).  Fills the lattice's offsets and returns the seed kind.
int aux_lattice(SeedLattice& lat, const float origin[3], const float lattice_spacing[3],
                const uint64_t shape[3], const float delta[3], uint64_t* particles);
// The per-point reduction of an auxiliary lattice's final positions (layout of
// seed kind `kind` over lat's d0 x d1 x d2 points) into out, on stream s.
void aux_reduce(const float* finals, int kind, const SeedLattice& lat, const float delta[3],
                float step_size, uint64_t iterations, float* out, hipStream_t s);

// Auxiliary-neighbour FTLE into a device volume; run_ftle_aux returns it in a
// result (the serving engine's per-request form).
void ftle_aux_into(dpah_context* ctx, const dpah_field* field, const float origin[3],
                   const float lattice_spacing[3], const uint64_t lattice_shape[3],
                   const float delta[3], uint64_t iterations, float step_size, float start_time,
                   int integrator, float* ftle_out, bool persistent, uint64_t* total_steps);
dpah_result* run_ftle_aux(dpah_context* ctx, const dpah_field* field, const float origin[3],
                          const float lattice_spacing[3], const uint64_t lattice_shape[3],
                          const float delta[3], uint64_t iterations, float step_size,
                          float start_time, int integrator, bool persistent);

// Slot shift of a recorded persistent launch of `rows` rows: launches large
// enough to stream their slots to HBM stage whole blocks of eight slots in
// LDS; smaller ones (slots L2-resident, few latency-bound waves) store each
// slot as it comes (ABC 64^3, 4096 rows x 258 slots: 0.317 ms with slot
// shift 0, 0.356 ms with blocks of eight stored directly, 0.39 ms staged).
constexpr uint64_t kStageMinRows = 1ull << 15;
inline uint32_t slot_shift_for(uint64_t rows, bool persistent, int slot_stage = 0) {
  if (!persistent || slot_stage == 1 || slot_stage == 3) return 0u;
  return slot_stage == 2 || rows >= kStageMinRows ? kStagedShift : 0u;
}
// Whether a recorded launch runs one thread per row instead of the persistent
// kernel.  By default a single request's rounds do (dpah_advect: the largest
// 64-job batch job alone, 243,573 rows x 3679 steps, advect 14.0 -> 10.7 ms
// and compaction 4.3 -> 5.5 ms; 100k vortex rows x 1000: 2.17 -> 1.98 ms with
// the compaction; ABC 64^3: 0.243 -> 0.226 ms; served pathlines on a 256^3 x
// 16-slice field: 6.1 -> 5.2 ms), and a batch's launch over many jobs' rounds
// stays persistent with staged blocks (the 64-job batch: advect 24.5 ms
// against 29.1-30.1 one thread per row, compaction 18.6-19.1 against
// 24.2-24.5, the one-slot layout leaving 12-byte holes of retired rows).
inline bool recorded_grid(int slot_stage, bool batch) {
  return slot_stage == 3 || (slot_stage == 0 && !batch);
}
// The same under the tuning of recorded launches on field f with this
// integrator (an autotuned entry or the context's).
uint32_t record_slot_shift(const dpah_context* ctx, const FieldView& f, int integrator,
                           uint64_t rows, bool persistent);
// recorded_grid under the same tuning.
bool record_grid(const dpah_context* ctx, const FieldView& f, int integrator, bool batch);

// Prunes, indexes and colours `count` slot rows (slot shift g, slot_float);
// allocates the chunk's outputs.  Synchronises the stream twice (the output
// size, then the written chunk).
Chunk compact_rows(dpah_context* ctx, const float* slots, uint64_t stride, uint32_t g,
                   const uint32_t* steps, const uint8_t* special, uint64_t count,
                   uint64_t vertex_base, int use64, hipStream_t stream = nullptr);

// The same in two phases, for callers that compact many row sets with one
// synchronisation: compact_count launches the per-row counts and their scan
// (vcount / icount: `count` words each; sums: compact_sums_words(count) words,
// whose last two hold the vertex and segment totals once the stream reaches
// them); compact_write allocates the chunk for those totals and launches the
// write.  Neither waits for the device.
struct CompactRows {
  const float* slots;
  uint64_t stride;
  const uint32_t* steps;
  const uint8_t* special;
  uint64_t count;
  uint64_t vertex_base;
  int use64;
  uint32_t slot_shift;
};
uint64_t compact_sums_words(uint64_t count);
void compact_count(dpah_context* ctx, const CompactRows& r, uint64_t* vcount, uint64_t* icount,
                   uint64_t* sums, hipStream_t stream);
Chunk compact_write(dpah_context* ctx, const CompactRows& r, const uint64_t* vcount,
                    const uint64_t* icount, uint64_t vertices, uint64_t segments,
                    hipStream_t stream);
// The writes of several rounds (counted and scanned, totals read back) in one
// launch: the chunks in job order (empty for empty rounds).  `table` receives
// the launch's segment table, which must outlive the launch (keep it until
// the stream is synchronised).  On an error no chunk is left allocated.
struct WriteJob {
  CompactRows rows;
  const uint64_t* vcount;
  const uint64_t* icount;
  uint64_t vertices, segments;
};
std::vector<Chunk> compact_write_many(dpah_context* ctx, const std::vector<WriteJob>& jobs,
                                      DeviceBuffer& table, hipStream_t stream);

}  // namespace dpah
