// serve.hip -- the serving engine: resident fields, memory-based admission,
// continuous batching of advection and FTLE queries (serve_kernels.hpp).
//
// One engine thread owns the device.  Between ticks it admits queued requests
// whose memory footprint fits the budget (FIFO, no overtaking), initialises
// their particle state, and retires finished ones (compaction / FTLE reduce,
// then the ticket becomes ready).  Each tick launches one kernel per
// (field, integrator, record) group that advances every admitted particle by
// up to `steps_per_tick` steps; with no step budget an FTLE request instead
// runs as its own persistent lattice launch (the brick-walked queue of
// dpah_ftle_aux_device).  Up to three launches run at once, one per stream.
// Two comparators run each request alone, in arrival order: the per-request
// mode allocates, uploads, launches one thread per particle, compacts and
// frees for every request; the serial mode goes through the library's
// single-request path (dpah_advect / dpah_ftle_aux_device on the context's one
// stream).  Results stay on the device, or are copied to host memory in every
// mode alike (results_on_host).
#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <vector>

#include "ftle_kernels.hpp"
#include "runtime.hpp"
#include "serve_kernels.hpp"

using namespace dpah;

namespace {

// A new request's device counters: particles not yet retired, particle-steps.
__global__ void k_request_counters(unsigned int* left, unsigned long long* steps, unsigned int n) {
  if (threadIdx.x == 0) {
    *left = n;
    *steps = 0;
  }
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ReqState { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3, kCancelled = 4 };

struct RoundState {
  uint64_t first = 0, count = 0, stride = 0;
  DeviceBuffer slots;
};

struct Request {
  uint64_t ticket = 0;
  dpah_request spec{};
  std::vector<float> seeds;
  std::vector<uint64_t> rem;
  uint64_t count = 0;           // particles (6 per lattice point for FTLE)
  uint64_t footprint = 0;
  int state = kPending;
  int error = DPAH_OK;
  std::string message;
  bool cancel = false;
  double t_submit = 0, t_admit = 0, t_done = 0;
  double t_launch = 0, t_device = 0;   // first launch enqueued, advection seen done
  uint64_t steps = 0;
  // device state while running
  int slot = -1;
  DeviceBuffer pos, rem_d, it, status, special;
  DeviceBuffer seeds_d, rem_in_d;      // staged inputs
  std::vector<RoundState> rounds;
  bool set_up = false;
  unsigned int left_h = 0;             // host staging of the device counters
  unsigned long long steps_h = 0;
  dpah_result* result = nullptr;
};

uint64_t ftle_points(const dpah_request& r) {
  return r.lattice_shape[0] * r.lattice_shape[1] * r.lattice_shape[2];
}

template <bool kRecord, int kIntegrator, bool kUnsteady, int kDiv>
void launch_tick_t(dpah_context* ctx, const TickArgs& a, hipStream_t s) {
  const uint64_t grid = (a.n + 255) / 256;
  DPAH_REQUIRE(grid < (1ull << 31), "too many particles in one launch");
  ctx->timed("advect_tick", s, [&] {
    hipLaunchKernelGGL((k_advect_tick<kRecord, kIntegrator, kUnsteady, kDiv>), dim3((unsigned)grid),
                       dim3(256), 0, s, a);
  });
  DPAH_HIP(hipGetLastError());
}

template <bool kRecord, int kIntegrator, bool kUnsteady>
void launch_tick_div(dpah_context* ctx, const TickArgs& a, hipStream_t s) {
  switch (div_mode(a.f.exactdiv, a.f.fastdiv, kUnsteady)) {
    case kDivExact: launch_tick_t<kRecord, kIntegrator, kUnsteady, kDivExact>(ctx, a, s); break;
    case kDivFast:  launch_tick_t<kRecord, kIntegrator, kUnsteady, kDivFast>(ctx, a, s); break;
    default:        launch_tick_t<kRecord, kIntegrator, kUnsteady, kDivide>(ctx, a, s); break;
  }
}

void launch_tick(dpah_context* ctx, const TickArgs& a, int integrator, bool record,
                 hipStream_t s) {
  if (a.n == 0) return;
  const bool unsteady = a.f.nt > 1;
  if (record) {
    if (integrator == kRk4) {
      if (unsteady) launch_tick_div<true, kRk4, true>(ctx, a, s); else launch_tick_div<true, kRk4, false>(ctx, a, s);
    } else {
      if (unsteady) launch_tick_div<true, kDpaRk4, true>(ctx, a, s); else launch_tick_div<true, kDpaRk4, false>(ctx, a, s);
    }
  } else {
    if (integrator == kRk4) {
      if (unsteady) launch_tick_div<false, kRk4, true>(ctx, a, s); else launch_tick_div<false, kRk4, false>(ctx, a, s);
    } else {
      if (unsteady) launch_tick_div<false, kDpaRk4, true>(ctx, a, s); else launch_tick_div<false, kDpaRk4, false>(ctx, a, s);
    }
  }
}

}  // namespace

struct Launch {
  bool busy = false;
  bool ftle = false;   // one FTLE request as a persistent lattice launch (else a tick)
  std::vector<std::shared_ptr<Request>> reqs;
  DeviceBuffer segs;
  hipEvent_t done = nullptr;
  unsigned long long* counters = nullptr;   // this stream's step counter and queue heads
};

struct dpah_server {
  // Three streams: with the context's own stream that makes four, the number
  // of hardware queues a process gets by default (GPU_MAX_HW_QUEUES); a fifth
  // stream shares a queue with another and waits behind its work.  On four
  // resident 256^3 x 16 fields at 500 req/s: 4 streams 1.30e11 steps/s, p99
  // 234-247 ms; 3 streams 1.36e11, p99 106-114 ms; 2 streams 1.36e11, p50
  // 24 ms instead of 13.
  static constexpr int kStreams = 3;
  dpah_context* ctx = nullptr;
  dpah_server_options opt{};
  uint64_t budget = 0;
  uint64_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;   // submitted, not admitted
  std::vector<std::shared_ptr<Request>> ready;    // admitted, waiting for a stream
  std::map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread engine;
  // per admitted request (slot): particles not yet retired, particle-steps
  unsigned int* d_left = nullptr;
  unsigned long long* d_steps = nullptr;
  std::vector<int> free_slots;
  hipStream_t streams[kStreams] = {};
  Launch launches[kStreams];
  dpah_server_stats stats{};

  void run();
  void run_batched();
  void run_per_request();
  bool admit_locked();
  void setup(Request& r, hipStream_t s);
  bool persistent_ftle(const Request& r) const;
  void launch_ftle(const std::shared_ptr<Request>& rp, int sid);
  bool launch_ready();
  bool poll_launches();
  void finish_all(std::vector<std::shared_ptr<Request>>& done, hipStream_t s);
  void finish_ftle(Launch& L, hipStream_t s);
  void release_request(const std::shared_ptr<Request>& r, int state, int err,
                       const std::string& msg);
  void complete(const std::shared_ptr<Request>& r, int state, int err, const std::string& msg);
};

namespace {

// Device bytes a request holds while it runs, plus an upper bound of its
// result (compacted curves or FTLE volume).
uint64_t footprint(const Request& r) {
  const uint64_t n = r.count;
  uint64_t b = n * (12 + 8 + 4 + 1 + 1) + 64;
  if (r.spec.kind == DPAH_REQUEST_FTLE_AUX) return b + ftle_points(r.spec) * 4;
  b += n * (12 + 8);   // staged seeds and remaining
  if (r.spec.params.record) {
    const uint64_t ib = r.spec.params.use_64_bit_indices ? 8 : 4;
    uint64_t verts = 0;
    for (uint64_t i = 0; i < n; ++i) verts += r.rem[i] + 2;
    const uint64_t ppr = r.spec.params.particles_per_round;
    uint64_t slots = 0, active = n;
    do {
      const uint64_t rc = std::min(ppr, active), first = active - rc;
      uint64_t mx = 0;
      for (uint64_t i = 0; i < rc; ++i) mx = std::max(mx, r.rem[first + i]);
      slots += slot_region_floats(rc, mx + 2, 0) * sizeof(float);
      active = first;
    } while (active > 0);
    b += slots + verts * 16 + verts * 2 * ib;
  }
  return b;
}

}  // namespace

void dpah_server::complete(const std::shared_ptr<Request>& r, int state, int err,
                           const std::string& msg) {
  std::lock_guard<std::mutex> lk(mu);
  r->state = state;
  r->error = err;
  r->message = msg;
  r->t_done = now_s();
  if (state == kDone) {
    stats.completed += 1;
    stats.particle_steps += r->steps;
  } else if (state == kCancelled) {
    stats.cancelled += 1;
  } else {
    stats.failed += 1;
  }
  cv_done.notify_all();
}

// Frees a request's running state and its admission slot, then completes it.
void dpah_server::release_request(const std::shared_ptr<Request>& r, int state, int err,
                                  const std::string& msg) {
  r->pos.reset(); r->rem_d.reset(); r->it.reset(); r->status.reset(); r->special.reset();
  r->seeds_d.reset(); r->rem_in_d.reset();
  r->rounds.clear();
  {
    std::lock_guard<std::mutex> lk(mu);
    in_use -= r->footprint;
    if (r->slot >= 0) free_slots.push_back(r->slot);
    r->slot = -1;
  }
  complete(r, state, err, msg);
}

// FIFO admission by memory footprint and admission slots (no overtaking).
bool dpah_server::admit_locked() {
  bool any = false;
  while (!pending.empty()) {
    auto r = pending.front();
    if (r->cancel) {
      r->state = kCancelled;
      r->t_done = now_s();
      stats.cancelled += 1;
      cv_done.notify_all();
      pending.pop_front();
      continue;
    }
    if (free_slots.empty() || in_use + r->footprint > budget) break;
    // and beside everything else the context holds, the finished results the
    // caller has not freed yet included (admission is FIFO, so an admitted
    // request always finishes and frees its place; a request larger than the
    // context's budget goes ahead when nothing else runs and fails alone)
    if (ctx->bytes_in_use.load() + r->footprint > ctx->memory_budget &&
        free_slots.size() < opt.max_active)
      break;
    r->slot = free_slots.back();
    free_slots.pop_back();
    in_use += r->footprint;
    r->state = kRunning;
    r->t_admit = now_s();
    ready.push_back(r);
    pending.pop_front();
    any = true;
  }
  return any;
}

// Device state of a newly admitted request, written on stream s.
void dpah_server::setup(Request& r, hipStream_t s) {
  const uint64_t n = r.count;
  const uint64_t n1 = std::max<uint64_t>(n, 1);
  r.pos = DeviceBuffer(ctx, 3 * n1 * sizeof(float));
  r.rem_d = DeviceBuffer(ctx, n1 * sizeof(uint64_t));
  r.it = DeviceBuffer(ctx, n1 * sizeof(uint32_t));
  r.status = DeviceBuffer(ctx, n1);
  r.special = DeviceBuffer(ctx, n1);
  const bool ftle = r.spec.kind == DPAH_REQUEST_FTLE_AUX;
  const bool record = !ftle && r.spec.params.record;
  TickInitArgs ia{};
  ia.count = n;
  ia.pos = r.pos.as<float>();
  ia.rem = r.rem_d.as<uint64_t>();
  ia.it = r.it.as<uint32_t>();
  ia.status = r.status.as<uint8_t>();
  ia.special = r.special.as<uint8_t>();
  if (ftle) {
    ia.seed_kind = aux_lattice(ia.lat, r.spec.origin, r.spec.lattice_spacing,
                               r.spec.lattice_shape, r.spec.delta, nullptr);
    ia.uniform_rem = r.spec.iterations;
  } else {
    // staged inputs live until the request finishes (another stream may reuse
    // a released block before this stream has read it)
    r.seeds_d = DeviceBuffer(ctx, 3 * n1 * sizeof(float));
    r.rem_in_d = DeviceBuffer(ctx, n1 * sizeof(uint64_t));
    if (n) {
      DPAH_HIP(hipMemcpyAsync(r.seeds_d.ptr, r.seeds.data(), 3 * n * sizeof(float),
                              hipMemcpyHostToDevice, s));
      DPAH_HIP(hipMemcpyAsync(r.rem_in_d.ptr, r.rem.data(), n * sizeof(uint64_t),
                              hipMemcpyHostToDevice, s));
    }
    ia.seed_kind = kSeedArray;
    ia.seeds = r.seeds_d.as<float>();
    ia.rem_in = r.rem_in_d.as<uint64_t>();
  }
  // dpa's rounds (tail first); each round has its own slot rows
  const uint64_t ppr = ftle ? std::max<uint64_t>(n, 1) : r.spec.params.particles_per_round;
  uint64_t active = n;
  do {
    RoundState rs;
    rs.count = std::min(ppr, active);
    rs.first = active - rs.count;
    uint64_t mx = 0;
    for (uint64_t i = 0; i < rs.count && !ftle; ++i) mx = std::max(mx, r.rem[rs.first + i]);
    rs.stride = mx + 2;
    if (record && rs.count)
      rs.slots = DeviceBuffer(ctx, slot_region_floats(rs.count, rs.stride, 0) * sizeof(float));
    r.rounds.push_back(std::move(rs));
    active = r.rounds.back().first;
  } while (active > 0);
  for (RoundState& rs : r.rounds) {
    if (!rs.count) continue;
    TickInitArgs a = ia;
    a.count = rs.count;
    if (!ftle) { a.seeds = ia.seeds + 3 * rs.first; a.rem_in = ia.rem_in + rs.first; }
    a.pos = ia.pos + 3 * rs.first;
    a.rem = ia.rem + rs.first;
    a.it = ia.it + rs.first;
    a.status = ia.status + rs.first;
    a.special = ia.special + rs.first;
    a.slots = record ? rs.slots.as<float>() : nullptr;
    a.stride = rs.stride;
    hipLaunchKernelGGL(k_tick_init, dim3((unsigned)((rs.count + 255) / 256)), dim3(256), 0, s, a);
  }
  DPAH_HIP(hipGetLastError());
  r.left_h = (unsigned int)n;
  r.steps_h = 0;
  hipLaunchKernelGGL(k_request_counters, dim3(1), dim3(64), 0, s, d_left + r.slot,
                     d_steps + r.slot, (unsigned int)n);
  DPAH_HIP(hipGetLastError());
  r.set_up = true;
}

// Starts one launch per free stream, each over the ready requests that share
// field, integrator and record flag (everything admitted since the last
// launch is coalesced into it).
bool dpah_server::launch_ready() {
  bool launched = false;
  while (!ready.empty()) {
    int sid = -1;
    for (int i = 0; i < kStreams; ++i)
      if (!launches[i].busy) { sid = i; break; }
    if (sid < 0) break;
    // one persistent FTLE launch at a time: each is sized to fill the chip on
    // its own, and several at once share the L2 between distant lattices;
    // the other streams take the next tick of pathlines meanwhile
    bool ftle_busy = false;
    for (int i = 0; i < kStreams; ++i) ftle_busy |= launches[i].busy && launches[i].ftle;
    size_t pick = ready.size();
    for (size_t i = 0; i < ready.size(); ++i)
      if (!persistent_ftle(*ready[i]) || !ftle_busy) { pick = i; break; }
    if (pick == ready.size()) break;
    if (persistent_ftle(*ready[pick])) {
      auto rp = ready[pick];
      ready.erase(ready.begin() + (std::ptrdiff_t)pick);
      launch_ftle(rp, sid);
      launched = true;
      continue;
    }
    const Request& head = *ready[pick];
    const bool rec0 = head.spec.kind == DPAH_REQUEST_ADVECT && head.spec.params.record;
    const auto key = std::make_tuple(head.spec.field, head.spec.params.integrator, rec0);
    Launch& L = launches[sid];
    L.ftle = false;
    L.reqs.clear();
    std::vector<std::shared_ptr<Request>> rest;
    for (auto& r : ready) {
      const bool rec = r->spec.kind == DPAH_REQUEST_ADVECT && r->spec.params.record;
      if (!persistent_ftle(*r) &&
          std::make_tuple(r->spec.field, r->spec.params.integrator, rec) == key)
        L.reqs.push_back(r);
      else
        rest.push_back(r);
    }
    ready.swap(rest);
    hipStream_t s = streams[sid];
    std::vector<TickSeg> segs;
    std::vector<std::shared_ptr<Request>> live;
    for (auto& rp : L.reqs) {
      Request& r = *rp;
      bool cancelled;
      {
        std::lock_guard<std::mutex> lk(mu);
        cancelled = r.cancel;
      }
      if (cancelled) {
        release_request(rp, kCancelled, DPAH_ERR_CANCELLED, "");
        continue;
      }
      try {
        if (!r.set_up) setup(r, s);
      } catch (const Error& e) {
        // setup may have queued copies into the request's buffers on s before
        // it failed: they finish before the buffers go back to the shared
        // cache, which the other streams draw from
        (void)hipStreamSynchronize(s);
        release_request(rp, kFailed, e.code, e.what());
        continue;
      }
      live.push_back(rp);
      if (r.t_launch == 0) r.t_launch = now_s();
      for (RoundState& rs : r.rounds) {
        if (!rs.count) continue;
        TickSeg t{};
        t.count = rs.count;
        t.pos = r.pos.as<float>() + 3 * rs.first;
        t.rem = r.rem_d.as<uint64_t>() + rs.first;
        t.it = r.it.as<uint32_t>() + rs.first;
        t.status = r.status.as<uint8_t>() + rs.first;
        t.special = r.special.as<uint8_t>() + rs.first;
        t.slots = rec0 ? rs.slots.as<float>() : nullptr;
        t.stride = rs.stride;
        t.dt = r.spec.params.step_size;
        t.t0 = r.spec.params.start_time;
        t.active_left = d_left + r.slot;
        t.steps = d_steps + r.slot;
        segs.push_back(t);
      }
    }
    L.reqs.swap(live);
    if (L.reqs.empty()) continue;
    uint64_t q = 0;   // segments start at multiples of 64 (wave-uniform in the kernel)
    for (auto& t : segs) { t.qbase = q; q += (t.count + 63) / 64 * 64; }
    L.segs = DeviceBuffer(ctx, std::max<size_t>(segs.size(), 1) * sizeof(TickSeg));
    if (!segs.empty())
      DPAH_HIP(hipMemcpyAsync(L.segs.ptr, segs.data(), segs.size() * sizeof(TickSeg),
                              hipMemcpyHostToDevice, s));
    TickArgs a{};
    a.f = std::get<0>(key)->view;
    bool scattered = true;   // FTLE requests seed lattices
    for (auto& rp : L.reqs) scattered &= rp->spec.kind == DPAH_REQUEST_ADVECT;
    if (!use_cells(ctx->tune, scattered)) a.f.cells = nullptr;
    if (!ctx->tune.fastdiv) a.f.fastdiv = a.f.exactdiv = 0;
    a.segs = L.segs.as<TickSeg>();
    a.nseg = (uint32_t)segs.size();
    a.n = q;
    a.budget = opt.steps_per_tick ? opt.steps_per_tick : 0xffffffffu;
    launch_tick(ctx, a, std::get<1>(key), rec0, s);
    DPAH_HIP(hipEventRecord(L.done, s));
    L.busy = true;
    launched = true;
    {
      std::lock_guard<std::mutex> lk(mu);
      stats.ticks += 1;
    }
  }
  return launched;
}

// Finishes the requests of every completed launch.
bool dpah_server::poll_launches() {
  bool progressed = false;
  for (int sid = 0; sid < kStreams; ++sid) {
    Launch& L = launches[sid];
    if (!L.busy) continue;
    const hipError_t q = hipEventQuery(L.done);
    if (q == hipErrorNotReady) continue;
    DPAH_HIP(q);
    hipStream_t s = streams[sid];
    if (L.ftle) {
      finish_ftle(L, s);
      L.reqs.clear();
      L.busy = false;
      progressed = true;
      continue;
    }
    // every slot's counter in one read (the engine thread holds the context,
    // so its pinned staging is free)
    auto* left = static_cast<unsigned int*>(ctx->host_staging(opt.max_active * sizeof(unsigned int)));
    DPAH_HIP(hipMemcpyAsync(left, d_left, opt.max_active * sizeof(unsigned int),
                            hipMemcpyDeviceToHost, s));
    DPAH_HIP(hipStreamSynchronize(s));
    for (auto& rp : L.reqs) rp->left_h = left[rp->slot];
    ctx->resolve_timings();
    const double t_seen = now_s();
    std::vector<std::shared_ptr<Request>> finished;
    for (auto& rp : L.reqs) {
      Request& r = *rp;
      r.t_device = t_seen;
      bool cancelled;
      {
        std::lock_guard<std::mutex> lk(mu);
        cancelled = r.cancel;
      }
      if (cancelled) {
        release_request(rp, kCancelled, DPAH_ERR_CANCELLED, "");
      } else if (r.left_h == 0) {
        finished.push_back(rp);
      } else {
        ready.push_back(rp);   // step budget used up: continue in a later launch
      }
    }
    finish_all(finished, s);
    for (auto& rp : finished) {
      if (rp->result) release_request(rp, kDone, DPAH_OK, "");
      else release_request(rp, kFailed, rp->error ? rp->error : DPAH_ERR_HIP, rp->message);
    }
    L.reqs.clear();
    L.segs.reset();
    L.busy = false;
    progressed = true;
  }
  return progressed;
}

// The results of every finished request of one launch, in two phases with one
// synchronisation each: first every request's step count, final states (or
// FTLE reduction) and every recorded round's vertex / index counts and scan,
// then every round's writes, sized by the totals.  A request that fails keeps
// no result and carries its error.
void dpah_server::finish_all(std::vector<std::shared_ptr<Request>>& done, hipStream_t s) {
  struct RoundCounts {
    size_t req = 0, round = 0;
    DeviceBuffer vcount, icount, sums;
    uint64_t totals[2] = {0, 0};
  };
  std::vector<std::unique_ptr<dpah_result>> res(done.size());
  std::vector<RoundCounts> counts;
  auto fail = [&](size_t i, const Error& e) {
    res[i].reset();
    done[i]->error = e.code;
    done[i]->message = e.what();
  };
  auto rows_of = [&](const Request& r, const RoundState& rs) {
    return CompactRows{rs.slots.as<float>(), rs.stride, r.it.as<uint32_t>() + rs.first,
                       r.special.as<uint8_t>() + rs.first, rs.count, 0,
                       r.spec.params.use_64_bit_indices ? 1 : 0, 0u};
  };
  // phase one's reads go to the context's pinned staging (the engine thread
  // holds the context): every slot's step count, then each advection
  // request's [remaining u64[n] | positions f32[3n] | status u8[n]]; requests
  // past the staging's limit read into their result vectors directly
  const uint64_t steps_bytes = (uint64_t)opt.max_active * sizeof(unsigned long long);
  std::vector<uint64_t> stage_off(done.size(), 0);
  uint64_t stage_bytes = steps_bytes;
  for (size_t i = 0; i < done.size(); ++i) {
    if (done[i]->spec.kind != DPAH_REQUEST_ADVECT) continue;
    stage_off[i] = stage_bytes;
    stage_bytes += (21 * done[i]->count + 7) / 8 * 8;
  }
  uint8_t* stage = static_cast<uint8_t*>(ctx->host_staging(stage_bytes));
  const bool staged = stage != nullptr;
  if (!staged) stage = static_cast<uint8_t*>(ctx->host_staging(steps_bytes));
  DPAH_HIP(hipMemcpyAsync(stage, d_steps, steps_bytes, hipMemcpyDeviceToHost, s));
  for (size_t i = 0; i < done.size(); ++i) {
    Request& r = *done[i];
    try {
      const bool ftle = r.spec.kind == DPAH_REQUEST_FTLE_AUX;
      res[i] = std::make_unique<dpah_result>();
      dpah_result& o = *res[i];
      o.ctx = ctx;
      const uint64_t n = r.count;
      if (ftle) {
        const uint64_t P = ftle_points(r.spec);
        o.ftle_count = P;
        o.d_ftle = static_cast<float*>(ctx->alloc(std::max<uint64_t>(P, 1) * sizeof(float)));
        SeedLattice lat;
        const int kind = aux_lattice(lat, r.spec.origin, r.spec.lattice_spacing,
                                     r.spec.lattice_shape, r.spec.delta, nullptr);
        aux_reduce(r.pos.as<float>(), kind, lat, r.spec.delta, r.spec.params.step_size,
                   r.spec.iterations, o.d_ftle, s);
        continue;
      }
      o.use64 = r.spec.params.use_64_bit_indices ? 1 : 0;
      o.final_pos.resize(3 * n);
      o.final_rem.resize(n);
      o.final_status.resize(n);
      if (n) {
        uint8_t* h = stage + stage_off[i];
        DPAH_HIP(hipMemcpyAsync(staged ? static_cast<void*>(h + 8 * n) : o.final_pos.data(),
                                r.pos.ptr, 3 * n * sizeof(float), hipMemcpyDeviceToHost, s));
        DPAH_HIP(hipMemcpyAsync(staged ? static_cast<void*>(h) : o.final_rem.data(), r.rem_d.ptr,
                                n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        DPAH_HIP(hipMemcpyAsync(staged ? static_cast<void*>(h + 20 * n) : o.final_status.data(),
                                r.status.ptr, n, hipMemcpyDeviceToHost, s));
      }
      if (!r.spec.params.record) continue;
      for (size_t k = 0; k < r.rounds.size(); ++k) {
        const RoundState& rs = r.rounds[k];
        if (!rs.count) continue;
        RoundCounts c;
        c.req = i;
        c.round = k;
        c.vcount = DeviceBuffer(ctx, rs.count * sizeof(uint64_t));
        c.icount = DeviceBuffer(ctx, rs.count * sizeof(uint64_t));
        const uint64_t words = compact_sums_words(rs.count);
        c.sums = DeviceBuffer(ctx, words * sizeof(uint64_t));
        compact_count(ctx, rows_of(r, rs), c.vcount.as<uint64_t>(), c.icount.as<uint64_t>(),
                      c.sums.as<uint64_t>(), s);
        counts.push_back(std::move(c));
      }
    } catch (const Error& e) {
      fail(i, e);
    }
  }
  for (RoundCounts& c : counts) {   // the totals sit in the last two words of each scan
    const uint64_t words = compact_sums_words(done[c.req]->rounds[c.round].count);
    DPAH_HIP(hipMemcpyAsync(c.totals, c.sums.as<uint64_t>() + words - 2, 2 * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
  }
  DPAH_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < done.size(); ++i) {
    Request& r = *done[i];
    r.steps_h = reinterpret_cast<const unsigned long long*>(stage)[r.slot];
    const uint64_t n = r.count;
    if (!res[i] || !staged || r.spec.kind != DPAH_REQUEST_ADVECT || !n) continue;
    const uint8_t* h = stage + stage_off[i];
    std::memcpy(res[i]->final_rem.data(), h, 8 * n);
    std::memcpy(res[i]->final_pos.data(), h + 8 * n, 12 * n);
    std::memcpy(res[i]->final_status.data(), h + 20 * n, n);
  }
  // every round of a request, empty ones too, in the request's round order
  for (size_t i = 0; i < done.size(); ++i)
    if (res[i] && done[i]->spec.kind == DPAH_REQUEST_ADVECT && done[i]->spec.params.record)
      res[i]->rounds.resize(done[i]->rounds.size());
  // every round's writes in one launch; a round that cannot be written
  // fails its request, and a launch that cannot be made fails them all
  std::vector<WriteJob> writes;
  std::vector<const RoundCounts*> written;
  for (RoundCounts& c : counts) {
    if (!res[c.req]) continue;
    if (!res[c.req]->use64 && c.totals[0] > 0xffffffffull) {
      fail(c.req, Error(DPAH_ERR_INVALID, "round has more than 2^32-1 vertices"));
      continue;
    }
    Request& r = *done[c.req];
    writes.push_back(WriteJob{rows_of(r, r.rounds[c.round]), c.vcount.as<uint64_t>(),
                              c.icount.as<uint64_t>(), c.totals[0], c.totals[1]});
    written.push_back(&c);
  }
  DeviceBuffer table;
  std::vector<Chunk> orphans;   // written for a request that failed meanwhile
  try {
    std::vector<Chunk> chunks = compact_write_many(ctx, writes, table, s);
    for (size_t k = 0; k < written.size(); ++k) {
      const RoundCounts& c = *written[k];
      if (!res[c.req]) {
        orphans.push_back(chunks[k]);
        continue;
      }
      Round& round = res[c.req]->rounds[c.round];
      round.V = chunks[k].V;
      round.I = chunks[k].I;
      round.chunks.push_back(chunks[k]);
    }
  } catch (const Error& e) {
    for (const RoundCounts* c : written)
      if (res[c->req]) fail(c->req, e);
  }
  DPAH_HIP(hipStreamSynchronize(s));   // the chunks are written before the counts are freed
  for (const Chunk& c : orphans)      // and before an orphan's block can be handed out again
    if (c.vertices) ctx->release(c.vertices, c.bytes);
  counts.clear();
  for (size_t i = 0; i < done.size(); ++i) {
    if (!res[i]) continue;
    Request& r = *done[i];
    res[i]->total_steps = r.steps_h;
    r.steps = r.steps_h;
    try {
      if (opt.results_on_host) materialize_host(res[i].get(), s);
      r.result = res[i].release();
    } catch (const Error& e) {
      fail(i, e);
    }
  }
}

// FTLE requests of a batched server that runs particles to retirement: each
// is one persistent launch of the auxiliary lattice (its queue walking 4x4x4
// bricks of points, as dpah_ftle_aux_device's does) on one of the server's
// streams, with that stream's own counters, then the per-point reduction.
bool dpah_server::persistent_ftle(const Request& r) const {
  if (r.spec.kind != DPAH_REQUEST_FTLE_AUX || opt.steps_per_tick != 0) return false;
  const uint64_t P = ftle_points(r.spec);
  return P > 0 && P <= (1ull << 25);
}

void dpah_server::launch_ftle(const std::shared_ptr<Request>& rp, int sid) {
  Request& r = *rp;
  bool cancelled;
  {
    std::lock_guard<std::mutex> lk(mu);
    cancelled = r.cancel;
  }
  if (cancelled) {
    release_request(rp, kCancelled, DPAH_ERR_CANCELLED, "");
    return;
  }
  Launch& L = launches[sid];
  hipStream_t s = streams[sid];
  const uint64_t P = ftle_points(r.spec);
  try {
    AdvectLaunch A{};
    A.f = &r.spec.field->view;
    A.seed_kind = aux_lattice(A.lat, r.spec.origin, r.spec.lattice_spacing, r.spec.lattice_shape,
                              r.spec.delta, &A.n);
    r.pos = DeviceBuffer(ctx, A.n * 3 * sizeof(float));   // the particles' final positions
    auto res = std::make_unique<dpah_result>();
    res->ctx = ctx;
    res->ftle_count = P;
    res->d_ftle = static_cast<float*>(ctx->alloc(P * sizeof(float)));
    A.uniform_rem = r.spec.iterations;
    A.pos_out = r.pos.as<float>();
    A.dt = r.spec.params.step_size;
    A.t0 = r.spec.params.start_time;
    A.integrator = r.spec.params.integrator;
    A.persistent = true;
    A.counters = L.counters;
    A.stream = s;
    launch_advect(ctx, A);
    aux_reduce(r.pos.as<float>(), A.seed_kind, A.lat, r.spec.delta, r.spec.params.step_size,
               r.spec.iterations, res->d_ftle, s);
    DPAH_HIP(hipEventRecord(L.done, s));
    r.result = res.release();
  } catch (const Error& e) {
    (void)hipStreamSynchronize(s);
    delete r.result;
    r.result = nullptr;
    release_request(rp, kFailed, e.code, e.what());
    return;
  }
  if (r.t_launch == 0) r.t_launch = now_s();
  L.ftle = true;
  L.reqs.assign(1, rp);
  L.busy = true;
  std::lock_guard<std::mutex> lk(mu);
  stats.ticks += 1;
}

void dpah_server::finish_ftle(Launch& L, hipStream_t s) {
  auto rp = L.reqs.front();
  Request& r = *rp;
  auto* steps = static_cast<unsigned long long*>(ctx->host_staging(sizeof(unsigned long long)));
  DPAH_HIP(hipMemcpyAsync(steps, L.counters + 1, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
  DPAH_HIP(hipStreamSynchronize(s));
  r.steps_h = *steps;
  ctx->resolve_timings();
  r.t_device = now_s();
  bool cancelled;
  {
    std::lock_guard<std::mutex> lk(mu);
    cancelled = r.cancel;
  }
  if (cancelled) {
    delete r.result;
    r.result = nullptr;
    release_request(rp, kCancelled, DPAH_ERR_CANCELLED, "");
    return;
  }
  r.steps = r.steps_h;
  r.result->total_steps = r.steps_h;
  try {
    if (opt.results_on_host) materialize_host(r.result, s);
    release_request(rp, kDone, DPAH_OK, "");
  } catch (const Error& e) {
    delete r.result;
    r.result = nullptr;
    release_request(rp, kFailed, e.code, e.what());
  }
}

void dpah_server::run_batched() {
  while (true) {
    {   // results the caller freed while a call held the context count as free again
      std::lock_guard<std::recursive_mutex> ctx_lk(ctx->mu);
      ctx->drain_deferred();
    }
    bool busy = false, free_stream = false;
    for (auto& L : launches) {
      busy |= L.busy;
      free_stream |= !L.busy;
    }
    {
      std::unique_lock<std::mutex> lk(mu);
      if (!busy && ready.empty())
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
      else if (busy && !stop && (ready.empty() || !free_stream))
        // launches in flight: poll their events every 50 us, or sooner when a
        // request arrives and a stream is free to take it
        cv_work.wait_for(lk, std::chrono::microseconds(50),
                         [&] { return stop || (free_stream && !pending.empty()); });
      if (stop) {
        for (auto& r : pending) r->cancel = true;
        for (auto& r : ready) r->cancel = true;
        for (auto& L : launches)
          for (auto& r : L.reqs) r->cancel = true;
        if (pending.empty() && ready.empty() && !busy) return;
      }
      admit_locked();
    }
    std::lock_guard<std::recursive_mutex> ctx_lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    const double t = now_s();
    const bool p = poll_launches();
    const bool l = launch_ready();
    if (p || l) {
      std::lock_guard<std::mutex> lk(mu);
      stats.busy_seconds += now_s() - t;
    }
  }
}

void dpah_server::run_per_request() {
  while (true) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
      if (pending.empty()) return;
      r = pending.front();
      pending.pop_front();
      if (r->cancel || stop) {
        r->state = kCancelled;
        r->t_done = now_s();
        stats.cancelled += 1;
        cv_done.notify_all();
        continue;
      }
    }
    // memory admission of the one-at-a-time modes: wait, without the
    // context's lock, until the request fits beside the results the caller
    // still holds (whose frees need that lock)
    // (a request larger than the whole budget goes ahead and fails on its own)
    while (r->footprint <= ctx->memory_budget &&
           ctx->bytes_in_use.load() + r->footprint > ctx->memory_budget) {
      {
        std::lock_guard<std::mutex> lk(mu);
        if (stop || r->cancel) break;
      }
      {
        std::lock_guard<std::recursive_mutex> ctx_lk(ctx->mu);
        ctx->drain_deferred();
      }
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      if (r->cancel || stop) {
        r->state = kCancelled;
        r->t_done = now_s();
        stats.cancelled += 1;
        cv_done.notify_all();
        continue;
      }
      r->state = kRunning;
      r->t_admit = now_s();
      r->t_launch = r->t_admit;
    }
    const double t = now_s();
    try {
      std::lock_guard<std::recursive_mutex> ctx_lk(ctx->mu);
      DPAH_HIP(hipSetDevice(ctx->device));
      // DPAH_SERVE_SERIAL: exactly what a caller gets from dpah_advect /
      // dpah_ftle_aux_device on this context (same kernels, ordering, caching
      // allocator).  DPAH_SERVE_PER_REQUEST: hipMalloc, upload, one thread per
      // particle, compaction, download and hipFree for every request.
      const bool naive = opt.mode == DPAH_SERVE_PER_REQUEST;
      struct CachingOff {
        dpah_context* c;
        bool was;
        CachingOff(dpah_context* x, bool on) : c(x), was(x->caching) { if (on) c->caching = false; }
        ~CachingOff() { c->caching = was; }
      } caching_off(ctx, naive);
      if (r->spec.kind == DPAH_REQUEST_FTLE_AUX) {
        r->result = run_ftle_aux(ctx, r->spec.field, r->spec.origin, r->spec.lattice_spacing,
                                 r->spec.lattice_shape, r->spec.delta, r->spec.iterations,
                                 r->spec.params.step_size, r->spec.params.start_time,
                                 r->spec.params.integrator, !naive);
      } else {
        r->result = run_advect(ctx, r->spec.field, r->seeds.data(), r->rem.data(), 0, r->count,
                               &r->spec.params, !naive);
      }
      if (opt.results_on_host) materialize_host(r->result, ctx->stream);
      r->steps = r->result->total_steps;
      r->t_device = now_s();
      complete(r, kDone, DPAH_OK, "");
    } catch (const Error& e) {
      complete(r, kFailed, e.code, e.what());
    }
    std::lock_guard<std::mutex> lk(mu);
    stats.busy_seconds += now_s() - t;
  }
}

void dpah_server::run() {
  if (opt.mode == DPAH_SERVE_PER_REQUEST || opt.mode == DPAH_SERVE_SERIAL) run_per_request();
  else run_batched();
}

// ------------------------------------------------------------------ ABI ----

DPAH_EXPORT int dpah_server_create(dpah_context* ctx, const dpah_server_options* options,
                                   dpah_server** out) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && out, "null argument");
    auto s = std::make_unique<dpah_server>();
    s->ctx = ctx;
    if (options) s->opt = *options;
    DPAH_REQUIRE(s->opt.mode == DPAH_SERVE_BATCHED || s->opt.mode == DPAH_SERVE_PER_REQUEST ||
                     s->opt.mode == DPAH_SERVE_SERIAL,
                 "unknown serving mode");
    if (!s->opt.max_active) s->opt.max_active = 256;
    // steps_per_tick 0: every launch runs its particles to completion
    s->budget = s->opt.memory_budget_bytes ? s->opt.memory_budget_bytes : ctx->memory_budget / 2;
    {
      std::lock_guard<std::recursive_mutex> lk(ctx->mu);
      DPAH_HIP(hipSetDevice(ctx->device));
      s->d_left = static_cast<unsigned int*>(ctx->alloc(s->opt.max_active * sizeof(unsigned int)));
      s->d_steps = static_cast<unsigned long long*>(
          ctx->alloc(s->opt.max_active * sizeof(unsigned long long)));
      DPAH_HIP(hipMemsetAsync(s->d_left, 0, s->opt.max_active * sizeof(unsigned int), ctx->stream));
      for (int i = 0; i < dpah_server::kStreams; ++i) {
        DPAH_HIP(hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking));
        DPAH_HIP(hipEventCreateWithFlags(&s->launches[i].done, hipEventDisableTiming));
        s->launches[i].counters = static_cast<unsigned long long*>(
            ctx->alloc(kCounterCount * sizeof(unsigned long long)));
        DPAH_HIP(hipMemsetAsync(s->launches[i].counters, 0,
                                kCounterCount * sizeof(unsigned long long), ctx->stream));
      }
      ctx->sync();
    }
    for (int i = (int)s->opt.max_active - 1; i >= 0; --i) s->free_slots.push_back(i);
    dpah_server* raw = s.get();
    s->engine = std::thread([raw] {
      try {
        raw->run();
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(raw->mu);
        raw->stats.failed += 1;
        for (auto& kv : raw->all)
          if (kv.second->state == kPending || kv.second->state == kRunning) {
            kv.second->state = kFailed;
            kv.second->error = DPAH_ERR_HIP;
            kv.second->message = std::string("engine stopped: ") + e.what();
          }
        raw->cv_done.notify_all();
      }
    });
    *out = s.release();
  });
}

DPAH_EXPORT int dpah_server_submit(dpah_server* s, const dpah_request* req, uint64_t* ticket) {
  return guarded([&] {
    DPAH_REQUIRE(s && req && ticket, "null argument");
    DPAH_REQUIRE(req->field && req->field->ctx == s->ctx, "request field is null or foreign");
    DPAH_REQUIRE(req->params.integrator == DPAH_INTEGRATOR_DPA_RK4 ||
                     req->params.integrator == DPAH_INTEGRATOR_RK4,
                 "unknown integrator");
    auto r = std::make_shared<Request>();
    r->spec = *req;
    if (req->kind == DPAH_REQUEST_FTLE_AUX) {
      DPAH_REQUIRE(req->iterations < (1ull << 32) - 2, "iterations must be below 2^32 - 2");
      SeedLattice lat;
      aux_lattice(lat, req->origin, req->lattice_spacing, req->lattice_shape, req->delta,
                  &r->count);
      r->spec.params.record = 0;
      r->spec.seeds = nullptr;
      r->spec.remaining = nullptr;
    } else {
      DPAH_REQUIRE(req->kind == DPAH_REQUEST_ADVECT, "unknown request kind");
      DPAH_REQUIRE(req->count == 0 || req->seeds, "seeds are null");
      DPAH_REQUIRE(req->count == 0 || req->params.particles_per_round > 0,
                   "particles_per_round must be positive");
      r->count = req->count;
      r->seeds.assign(req->seeds, req->seeds + 3 * req->count);
      r->rem.resize(req->count);
      for (uint64_t i = 0; i < req->count; ++i) {
        r->rem[i] = req->remaining ? req->remaining[i] : req->uniform_remaining;
        DPAH_REQUIRE(r->rem[i] <= (1ull << 32) - 3, "remaining iterations must be below 2^32 - 2");
      }
      r->spec.seeds = nullptr;
      r->spec.remaining = nullptr;
    }
    r->footprint = footprint(*r);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->opt.mode == DPAH_SERVE_BATCHED && r->footprint > s->budget) {
      s->stats.rejected += 1;
      throw Error(DPAH_ERR_ADMISSION, "request needs " + std::to_string(r->footprint) +
                                          " bytes, more than the server budget " +
                                          std::to_string(s->budget));
    }
    DPAH_REQUIRE(!s->stop, "server is shutting down");
    r->ticket = s->next_ticket++;
    r->t_submit = now_s();
    s->all[r->ticket] = r;
    s->pending.push_back(r);
    s->stats.submitted += 1;
    *ticket = r->ticket;
    s->cv_work.notify_all();
  });
}

DPAH_EXPORT int dpah_server_poll(dpah_server* s, uint64_t ticket, int* state) {
  return guarded([&] {
    DPAH_REQUIRE(s && state, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    DPAH_REQUIRE(it != s->all.end(), "unknown ticket");
    *state = it->second->state;
  });
}

DPAH_EXPORT int dpah_server_wait(dpah_server* s, uint64_t ticket, int64_t timeout_ms,
                                 dpah_result** out) {
  int code = DPAH_OK;
  const int rc = guarded([&] {
    DPAH_REQUIRE(s && out, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    DPAH_REQUIRE(it != s->all.end(), "unknown ticket");
    auto r = it->second;
    auto ready = [&] { return r->state >= kDone; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
      throw Error(DPAH_ERR_TIMEOUT, "request not finished");
    if (r->state == kCancelled) throw Error(DPAH_ERR_CANCELLED, "request was cancelled");
    if (r->state == kFailed) throw Error(r->error ? r->error : DPAH_ERR_HIP, r->message);
    *out = r->result;
    r->result = nullptr;   // ownership moves to the caller
    (void)code;
  });
  return rc;
}

DPAH_EXPORT int dpah_server_request_info(dpah_server* s, uint64_t ticket, dpah_request_info* info) {
  return guarded([&] {
    DPAH_REQUIRE(s && info, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    DPAH_REQUIRE(it != s->all.end(), "unknown ticket");
    const Request& r = *it->second;
    info->state = r.state;
    info->queue_seconds = r.t_admit > 0 ? r.t_admit - r.t_submit : 0.0;
    info->latency_seconds = r.t_done > 0 ? r.t_done - r.t_submit : 0.0;
    info->particle_steps = r.steps;
    info->footprint_bytes = r.footprint;
    info->launch_seconds = r.t_launch > 0 ? r.t_launch - r.t_submit : 0.0;
    info->device_seconds = r.t_device > 0 ? r.t_device - r.t_submit : 0.0;
  });
}

DPAH_EXPORT int dpah_server_cancel(dpah_server* s, uint64_t ticket) {
  return guarded([&] {
    DPAH_REQUIRE(s, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    DPAH_REQUIRE(it != s->all.end(), "unknown ticket");
    it->second->cancel = true;
    s->cv_work.notify_all();
  });
}

DPAH_EXPORT int dpah_server_forget(dpah_server* s, uint64_t ticket) {
  return guarded([&] {
    DPAH_REQUIRE(s, "null argument");
    std::shared_ptr<Request> r;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->all.find(ticket);
      DPAH_REQUIRE(it != s->all.end(), "unknown ticket");
      DPAH_REQUIRE(it->second->state >= kDone, "request still queued or running");
      r = it->second;
      s->all.erase(it);
    }
    if (r->result) {
      std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
      delete r->result;
      r->result = nullptr;
    }
  });
}

DPAH_EXPORT int dpah_server_get_stats(dpah_server* s, dpah_server_stats* out) {
  return guarded([&] {
    DPAH_REQUIRE(s && out, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pending = s->pending.size();
    out->bytes_in_use = s->in_use;
  });
}

DPAH_EXPORT int dpah_server_destroy(dpah_server* s) {
  return guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv_work.notify_all();
    }
    if (s->engine.joinable()) s->engine.join();
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    (void)hipSetDevice(s->ctx->device);
    for (auto& kv : s->all)
      if (kv.second->result) delete kv.second->result;
    s->all.clear();
    for (int i = 0; i < dpah_server::kStreams; ++i) {
      if (s->streams[i]) (void)hipStreamSynchronize(s->streams[i]);
      s->launches[i].segs.reset();
      if (s->launches[i].counters)
        s->ctx->release(s->launches[i].counters, kCounterCount * sizeof(unsigned long long));
      if (s->launches[i].done) (void)hipEventDestroy(s->launches[i].done);
      if (s->streams[i]) (void)hipStreamDestroy(s->streams[i]);
    }
    s->ctx->release(s->d_left, s->opt.max_active * sizeof(unsigned int));
    s->ctx->release(s->d_steps, s->opt.max_active * sizeof(unsigned long long));
    delete s;
  });
}

DPAH_EXPORT int dpah_result_ftle(const dpah_result* r, float* out, uint64_t* count) {
  return guarded([&] {
    DPAH_REQUIRE(r && count, "null argument");
    *count = r->ftle_count;
    if (out && r->ftle_count && !r->d_ftle) {
      std::memcpy(out, r->host_ftle.get(), r->ftle_count * sizeof(float));
    } else if (out && r->ftle_count) {
      std::lock_guard<std::recursive_mutex> lk(r->ctx->mu);
      DPAH_HIP(hipSetDevice(r->ctx->device));
      r->ctx->download(out, r->d_ftle, r->ftle_count * sizeof(float), r->ctx->stream);
      r->ctx->sync();
    }
  });
}
