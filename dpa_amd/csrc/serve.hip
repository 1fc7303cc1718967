// serve.hip -- the serving engine: resident fields, memory-based admission,
// continuous batching of advection and FTLE queries (serve_kernels.hpp).
//
// One engine thread owns the device.  Between ticks it admits queued requests
// whose memory footprint fits the budget (FIFO, no overtaking), initialises
// their particle state, and retires finished ones (compaction / FTLE reduce,
// then the ticket becomes ready).  Each tick launches one kernel per
// (field, integrator, record) group that advances every admitted particle by
// up to `steps_per_tick` steps.  The per-request mode instead runs each request
// alone, in arrival order, with allocation, upload, one non-persistent launch,
// compaction and release per request: the comparator for batching.
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
  uint64_t steps = 0;
  // device state while running
  int slot = -1;
  DeviceBuffer pos, rem_d, it, status, special;
  std::vector<RoundState> rounds;
  dpah_result* result = nullptr;
};

uint64_t ftle_points(const dpah_request& r) {
  return r.lattice_shape[0] * r.lattice_shape[1] * r.lattice_shape[2];
}

template <bool kRecord, int kIntegrator, bool kUnsteady>
void launch_tick_t(dpah_context* ctx, const TickArgs& a) {
  const int block = 256;
  uint64_t grid = (uint64_t)ctx->cu_count * ctx->tune.blocks_per_cu;
  const uint64_t need = (a.n + block - 1) / block;
  if (grid > need) grid = need;
  ctx->timed("advect_tick", [&] {
    hipLaunchKernelGGL((k_advect_tick<kRecord, kIntegrator, kUnsteady>), dim3((unsigned)grid),
                       dim3(block), 0, ctx->stream, a);
  });
  DPAH_HIP(hipGetLastError());
}

__global__ void k_reset_queue(unsigned long long* q) {
  if (threadIdx.x == 0) *q = 0ull;
}

void launch_tick(dpah_context* ctx, const TickArgs& a, int integrator, bool record) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, ctx->stream, a.queue);
  const bool unsteady = a.f.nt > 1;
  if (record) {
    if (integrator == kRk4) {
      if (unsteady) launch_tick_t<true, kRk4, true>(ctx, a); else launch_tick_t<true, kRk4, false>(ctx, a);
    } else {
      if (unsteady) launch_tick_t<true, kDpaRk4, true>(ctx, a); else launch_tick_t<true, kDpaRk4, false>(ctx, a);
    }
  } else {
    if (integrator == kRk4) {
      if (unsteady) launch_tick_t<false, kRk4, true>(ctx, a); else launch_tick_t<false, kRk4, false>(ctx, a);
    } else {
      if (unsteady) launch_tick_t<false, kDpaRk4, true>(ctx, a); else launch_tick_t<false, kDpaRk4, false>(ctx, a);
    }
  }
}

}  // namespace

struct dpah_server {
  dpah_context* ctx = nullptr;
  dpah_server_options opt{};
  uint64_t budget = 0;
  uint64_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::vector<std::shared_ptr<Request>> active;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread engine;
  // per active slot: [0] active_left (u32 in a u64 cell), steps
  unsigned int* d_left = nullptr;
  unsigned long long* d_steps = nullptr;
  std::vector<int> free_slots;
  dpah_server_stats stats{};
  DeviceBuffer seg_buf;       // tick segment table, grown on demand

  void run();
  void run_batched();
  void run_per_request();
  bool admit_locked(std::vector<std::shared_ptr<Request>>& admitted);
  void setup(Request& r);
  void tick();
  void poll_and_finish();
  void finish(Request& r);
  void complete(const std::shared_ptr<Request>& r, int state, int err, const std::string& msg);
};

namespace {

// Device bytes a request holds while it runs, plus an upper bound of its
// result (compacted curves or FTLE volume).
uint64_t footprint(const Request& r) {
  const uint64_t n = r.count;
  uint64_t b = n * (12 + 8 + 4 + 1 + 1) + 64;
  if (r.spec.kind == DPAH_REQUEST_FTLE_AUX) return b + ftle_points(r.spec) * 4;
  if (r.spec.params.record) {
    const uint64_t ib = r.spec.params.use_64_bit_indices ? 8 : 4;
    uint64_t verts = 0;
    for (uint64_t i = 0; i < n; ++i) verts += r.rem[i] + 2;
    // slots rows are at most (max remaining + 2) per particle of a round
    const uint64_t ppr = r.spec.params.particles_per_round;
    uint64_t slots = 0, active = n;
    do {
      const uint64_t rc = std::min(ppr, active), first = active - rc;
      uint64_t mx = 0;
      for (uint64_t i = 0; i < rc; ++i) mx = std::max(mx, r.rem[first + i]);
      slots += rc * (mx + 2) * 12;
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

bool dpah_server::admit_locked(std::vector<std::shared_ptr<Request>>& admitted) {
  while (!pending.empty()) {
    auto& r = pending.front();
    if (r->cancel) {
      r->state = kCancelled;
      r->t_done = now_s();
      stats.cancelled += 1;
      cv_done.notify_all();
      pending.pop_front();
      continue;
    }
    if (free_slots.empty() || in_use + r->footprint > budget) break;
    r->slot = free_slots.back();
    free_slots.pop_back();
    in_use += r->footprint;
    r->state = kRunning;
    r->t_admit = now_s();
    admitted.push_back(r);
    pending.pop_front();
  }
  return !admitted.empty();
}

void dpah_server::setup(Request& r) {
  const uint64_t n = r.count;
  const uint64_t n1 = std::max<uint64_t>(n, 1);
  r.pos = DeviceBuffer(ctx, 3 * n1 * sizeof(float));
  r.rem_d = DeviceBuffer(ctx, n1 * sizeof(uint64_t));
  r.it = DeviceBuffer(ctx, n1 * sizeof(uint32_t));
  r.status = DeviceBuffer(ctx, n1);
  r.special = DeviceBuffer(ctx, n1);
  const bool ftle = r.spec.kind == DPAH_REQUEST_FTLE_AUX;
  const bool record = !ftle && r.spec.params.record;
  DeviceBuffer seeds, rem;
  TickInitArgs ia{};
  ia.count = n;
  ia.pos = r.pos.as<float>();
  ia.rem = r.rem_d.as<uint64_t>();
  ia.it = r.it.as<uint32_t>();
  ia.status = r.status.as<uint8_t>();
  ia.special = r.special.as<uint8_t>();
  if (ftle) {
    ia.seed_kind = kSeedAux;
    SeedLattice& L = ia.lat;
    L.lo0 = r.spec.origin[0]; L.lo1 = r.spec.origin[1]; L.lo2 = r.spec.origin[2];
    L.st0 = r.spec.lattice_spacing[0]; L.st1 = r.spec.lattice_spacing[1];
    L.st2 = r.spec.lattice_spacing[2];
    L.d0 = r.spec.lattice_shape[0]; L.d1 = r.spec.lattice_shape[1]; L.d2 = r.spec.lattice_shape[2];
    L.dl0 = r.spec.delta[0]; L.dl1 = r.spec.delta[1]; L.dl2 = r.spec.delta[2];
    ia.uniform_rem = r.spec.iterations;
  } else {
    seeds = DeviceBuffer(ctx, 3 * n1 * sizeof(float));
    rem = DeviceBuffer(ctx, n1 * sizeof(uint64_t));
    if (n) {
      DPAH_HIP(hipMemcpyAsync(seeds.ptr, r.seeds.data(), 3 * n * sizeof(float),
                              hipMemcpyHostToDevice, ctx->stream));
      DPAH_HIP(hipMemcpyAsync(rem.ptr, r.rem.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    ia.seed_kind = kSeedArray;
    ia.seeds = seeds.as<float>();
    ia.rem_in = rem.as<uint64_t>();
  }
  // rounds (tail first); each round has its own slot rows
  const uint64_t ppr = ftle ? std::max<uint64_t>(n, 1) : r.spec.params.particles_per_round;
  uint64_t active = n;
  do {
    RoundState rs;
    rs.count = std::min(ppr, active);
    rs.first = active - rs.count;
    uint64_t mx = 0;
    for (uint64_t i = 0; i < rs.count && !ftle; ++i) mx = std::max(mx, r.rem[rs.first + i]);
    rs.stride = mx + 2;
    if (record && rs.count) rs.slots = DeviceBuffer(ctx, rs.count * rs.stride * 3 * sizeof(float));
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
    hipLaunchKernelGGL(k_tick_init, dim3((unsigned)((rs.count + 255) / 256)), dim3(256), 0,
                       ctx->stream, a);
  }
  DPAH_HIP(hipGetLastError());
  const unsigned int left = (unsigned int)n;
  DPAH_HIP(hipMemcpyAsync(d_left + r.slot, &left, sizeof(left), hipMemcpyHostToDevice,
                          ctx->stream));
  const unsigned long long zero = 0;
  DPAH_HIP(hipMemcpyAsync(d_steps + r.slot, &zero, sizeof(zero), hipMemcpyHostToDevice,
                          ctx->stream));
  ctx->sync();   // seeds / rem staging buffers die here
}

void dpah_server::tick() {
  // group running requests by launch-uniform arguments
  std::map<std::tuple<const dpah_field*, int, int>, std::vector<TickSeg>> groups;
  for (auto& rp : active) {
    Request& r = *rp;
    const bool ftle = r.spec.kind == DPAH_REQUEST_FTLE_AUX;
    const bool record = !ftle && r.spec.params.record;
    const int integ = r.spec.params.integrator;
    auto& segs = groups[{r.spec.field, integ, record ? 1 : 0}];
    for (RoundState& rs : r.rounds) {
      if (!rs.count) continue;
      TickSeg s{};
      s.count = rs.count;
      s.pos = r.pos.as<float>() + 3 * rs.first;
      s.rem = r.rem_d.as<uint64_t>() + rs.first;
      s.it = r.it.as<uint32_t>() + rs.first;
      s.status = r.status.as<uint8_t>() + rs.first;
      s.special = r.special.as<uint8_t>() + rs.first;
      s.slots = record ? rs.slots.as<float>() : nullptr;
      s.stride = rs.stride;
      s.dt = r.spec.params.step_size;
      s.t0 = r.spec.params.start_time;
      s.active_left = d_left + r.slot;
      s.steps = d_steps + r.slot;
      segs.push_back(s);
    }
  }
  for (auto& g : groups) {
    auto& segs = g.second;
    if (segs.empty()) continue;
    uint64_t q = 0;
    for (auto& s : segs) { s.qbase = q; q += s.count; }
    const uint64_t need = segs.size() * sizeof(TickSeg);
    if (seg_buf.bytes < need) seg_buf = DeviceBuffer(ctx, std::max<uint64_t>(need, 64 * sizeof(TickSeg)));
    DPAH_HIP(hipMemcpyAsync(seg_buf.ptr, segs.data(), need, hipMemcpyHostToDevice, ctx->stream));
    TickArgs a{};
    a.f = std::get<0>(g.first)->view;
    a.segs = seg_buf.as<TickSeg>();
    a.nseg = (uint32_t)segs.size();
    a.n = q;
    a.budget = std::max<uint32_t>(1, opt.steps_per_tick);
    a.queue = ctx->d_counters + 3;
    launch_tick(ctx, a, std::get<1>(g.first), std::get<2>(g.first) != 0);
    ctx->sync();   // the table is rewritten by the next group
  }
  stats.ticks += 1;
}

void dpah_server::finish(Request& r) {
  const bool ftle = r.spec.kind == DPAH_REQUEST_FTLE_AUX;
  auto res = std::make_unique<dpah_result>();
  res->ctx = ctx;
  const uint64_t n = r.count;
  unsigned long long steps = 0;
  DPAH_HIP(hipMemcpyAsync(&steps, d_steps + r.slot, sizeof(steps), hipMemcpyDeviceToHost,
                          ctx->stream));
  if (ftle) {
    const uint64_t P = ftle_points(r.spec);
    res->ftle_count = P;
    res->d_ftle = static_cast<float*>(ctx->alloc(std::max<uint64_t>(P, 1) * sizeof(float)));
    FtleAuxArgs a{};
    a.finals = r.pos.as<float>();
    a.points = P;
    a.twod0 = 2.0f * r.spec.delta[0]; a.twod1 = 2.0f * r.spec.delta[1];
    a.twod2 = 2.0f * r.spec.delta[2];
    a.T_abs = std::fabs(r.spec.params.step_size * (float)r.spec.iterations);
    a.out = res->d_ftle;
    if (P)
      hipLaunchKernelGGL(k_ftle_aux_reduce, dim3((unsigned)((P + 255) / 256)), dim3(256), 0,
                         ctx->stream, a);
    DPAH_HIP(hipGetLastError());
  } else {
    res->use64 = r.spec.params.use_64_bit_indices ? 1 : 0;
    res->final_pos.resize(3 * n);
    res->final_rem.resize(n);
    res->final_status.resize(n);
    if (n) {
      DPAH_HIP(hipMemcpyAsync(res->final_pos.data(), r.pos.ptr, 3 * n * sizeof(float),
                              hipMemcpyDeviceToHost, ctx->stream));
      DPAH_HIP(hipMemcpyAsync(res->final_rem.data(), r.rem_d.ptr, n * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
      DPAH_HIP(hipMemcpyAsync(res->final_status.data(), r.status.ptr, n, hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    if (r.spec.params.record) {
      for (RoundState& rs : r.rounds) {
        Round round;
        if (rs.count) {
          Chunk ch = compact_rows(ctx, rs.slots.as<float>(), rs.stride,
                                  r.it.as<uint32_t>() + rs.first,
                                  r.special.as<uint8_t>() + rs.first, rs.count, 0, res->use64);
          round.V = ch.V;
          round.I = ch.I;
          round.chunks.push_back(ch);
        }
        res->rounds.push_back(std::move(round));
      }
    }
  }
  ctx->sync();
  res->total_steps = steps;
  r.steps = steps;
  r.result = res.release();
  // release the running state
  r.pos.reset(); r.rem_d.reset(); r.it.reset(); r.status.reset(); r.special.reset();
  r.rounds.clear();
}

void dpah_server::poll_and_finish() {
  if (active.empty()) return;
  std::vector<unsigned int> left(opt.max_active);
  DPAH_HIP(hipMemcpyAsync(left.data(), d_left, opt.max_active * sizeof(unsigned int),
                          hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  std::vector<std::shared_ptr<Request>> still;
  for (auto& rp : active) {
    Request& r = *rp;
    bool cancelled;
    {
      std::lock_guard<std::mutex> lk(mu);
      cancelled = r.cancel;
    }
    if (cancelled) {
      r.pos.reset(); r.rem_d.reset(); r.it.reset(); r.status.reset(); r.special.reset();
      r.rounds.clear();
    } else if (left[r.slot] == 0) {
      try {
        finish(r);
      } catch (const Error& e) {
        r.rounds.clear();
        complete(rp, kFailed, e.code, e.what());
        std::lock_guard<std::mutex> lk(mu);
        in_use -= r.footprint;
        free_slots.push_back(r.slot);
        continue;
      }
    } else {
      still.push_back(rp);
      continue;
    }
    complete(rp, cancelled ? kCancelled : kDone, cancelled ? DPAH_ERR_CANCELLED : DPAH_OK, "");
    std::lock_guard<std::mutex> lk(mu);
    in_use -= r.footprint;
    free_slots.push_back(r.slot);
  }
  active.swap(still);
}

void dpah_server::run_batched() {
  while (true) {
    std::vector<std::shared_ptr<Request>> admitted;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
      if (stop && pending.empty() && active.empty()) return;
      if (stop) {   // cancel what is left
        for (auto& r : pending) r->cancel = true;
        for (auto& r : active) r->cancel = true;
      }
      admit_locked(admitted);
    }
    std::lock_guard<std::recursive_mutex> ctx_lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    for (auto& r : admitted) {
      try {
        setup(*r);
        active.push_back(r);
      } catch (const Error& e) {
        r->rounds.clear();
        complete(r, kFailed, e.code, e.what());
        std::lock_guard<std::mutex> lk(mu);
        in_use -= r->footprint;
        free_slots.push_back(r->slot);
      }
    }
    if (!active.empty()) {
      const double t = now_s();
      tick();
      poll_and_finish();
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
      r->state = kRunning;
      r->t_admit = now_s();
    }
    const double t = now_s();
    try {
      std::lock_guard<std::recursive_mutex> ctx_lk(ctx->mu);
      DPAH_HIP(hipSetDevice(ctx->device));
      // the per-request form allocates and frees through the runtime every time
      struct Uncached {
        dpah_context* c;
        bool was;
        explicit Uncached(dpah_context* x) : c(x), was(x->caching) { c->caching = false; }
        ~Uncached() { c->caching = was; }
      } uncached(ctx);
      if (r->spec.kind == DPAH_REQUEST_FTLE_AUX) {
        r->result = run_ftle_aux(ctx, r->spec.field, r->spec.origin, r->spec.lattice_spacing,
                                 r->spec.lattice_shape, r->spec.delta, r->spec.iterations,
                                 r->spec.params.step_size, r->spec.params.start_time,
                                 r->spec.params.integrator, false);
      } else {
        r->result = run_advect(ctx, r->spec.field, r->seeds.data(), r->rem.data(), 0, r->count,
                               &r->spec.params, false);
      }
      r->steps = r->result->total_steps;
      complete(r, kDone, DPAH_OK, "");
    } catch (const Error& e) {
      complete(r, kFailed, e.code, e.what());
    }
    std::lock_guard<std::mutex> lk(mu);
    stats.busy_seconds += now_s() - t;
  }
}

void dpah_server::run() {
  if (opt.mode == DPAH_SERVE_PER_REQUEST) run_per_request();
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
    if (!s->opt.max_active) s->opt.max_active = 256;
    if (!s->opt.steps_per_tick) s->opt.steps_per_tick = 256;
    s->budget = s->opt.memory_budget_bytes ? s->opt.memory_budget_bytes : ctx->memory_budget / 2;
    {
      std::lock_guard<std::recursive_mutex> lk(ctx->mu);
      DPAH_HIP(hipSetDevice(ctx->device));
      s->d_left = static_cast<unsigned int*>(ctx->alloc(s->opt.max_active * sizeof(unsigned int)));
      s->d_steps = static_cast<unsigned long long*>(
          ctx->alloc(s->opt.max_active * sizeof(unsigned long long)));
      DPAH_HIP(hipMemsetAsync(s->d_left, 0, s->opt.max_active * sizeof(unsigned int), ctx->stream));
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
      r->count = 6 * ftle_points(*req);
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
    if (s->opt.mode != DPAH_SERVE_PER_REQUEST && r->footprint > s->budget) {
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
    s->seg_buf.reset();
    s->ctx->release(s->d_left, s->opt.max_active * sizeof(unsigned int));
    s->ctx->release(s->d_steps, s->opt.max_active * sizeof(unsigned long long));
    delete s;
  });
}

DPAH_EXPORT int dpah_result_ftle(const dpah_result* r, float* out, uint64_t* count) {
  return guarded([&] {
    DPAH_REQUIRE(r && count, "null argument");
    *count = r->ftle_count;
    if (out && r->ftle_count) {
      std::lock_guard<std::recursive_mutex> lk(r->ctx->mu);
      DPAH_HIP(hipSetDevice(r->ctx->device));
      DPAH_HIP(hipMemcpyAsync(out, r->d_ftle, r->ftle_count * sizeof(float), hipMemcpyDeviceToHost,
                              r->ctx->stream));
      r->ctx->sync();
    }
  });
}
