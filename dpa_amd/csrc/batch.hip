// batch.hip -- many seed sets of uneven size and length in one persistent
// launch.  Every job is one dpa rank's advection stage (dpah_advect); its
// rounds become segments of one on-device work queue, ordered longest first,
// so the launch replaces dpa's rank-level load balancing.  Each job's result is
// bit-identical to running it alone (particles are independent).
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "runtime.hpp"

using namespace dpah;

namespace {

struct RoundPlan {
  uint64_t first = 0;   // job-local first particle of the round
  uint64_t count = 0;
  uint64_t stride = 0;
  int seg = -1;         // segment index in the launch table, -1 if empty
};

struct JobPlan {
  uint64_t job = 0;     // index into the caller's array
  uint64_t base = 0;    // first particle in the group's concatenated arrays
  uint64_t count = 0;
  int use64 = 0;
  uint64_t rem_sum = 0;   // the job's starting iterations, summed
  std::vector<RoundPlan> rounds;
};

// The uniform starting iterations of jobs without a per-particle array.
__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void validate(const dpah_context* ctx, const dpah_job& j) {
  DPAH_REQUIRE(j.field, "job field is null");
  DPAH_REQUIRE(j.field->ctx == ctx, "job field belongs to another context");
  DPAH_REQUIRE(j.count == 0 || j.seeds, "job seeds are null");
  DPAH_REQUIRE(j.count == 0 || j.params.particles_per_round > 0,
               "particles_per_round must be positive");
  DPAH_REQUIRE(j.params.integrator == DPAH_INTEGRATOR_DPA_RK4 ||
                   j.params.integrator == DPAH_INTEGRATOR_RK4,
               "unknown integrator");
}

uint32_t float_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// One group of jobs that share field, step size, start time, integrator and
// record flag, run as one persistent launch followed by per-round compaction,
// in three stages:
//   plan     host planning, device buffers, seeds and segments uploaded;
//   enqueue  counter reset, the persistent launch and every recorded round's
//            vertex / index counts and scan: kernels only, no host
//            synchronisation and no allocation, so a stream capture can hold
//            it and replays recompute every output from the uploaded seeds;
//   finish   totals and final states read back in one synchronisation, then
//            every round's writes (sized by the totals) and the results.
struct GroupRun {
  bool record = false;
  uint64_t N = 0, q = 0, slot_floats = 0;
  uint32_t slot_shift = 0;
  std::vector<JobPlan> plans;
  std::vector<Segment> segs;
  std::vector<uint64_t> row_off, sums_off;
  std::vector<int> seg_use64;
  uint64_t rows_total = 0, sums_total = 0;
  AdvectLaunch L{};
  DeviceBuffer d_seeds, d_rem, d_pos, d_rem_out, d_status, d_steps, d_special, d_segs, d_slots;
  DeviceBuffer d_vcount, d_icount, d_sums;

  CompactRows rows_of(size_t s) const {
    return CompactRows{d_slots.as<float>() + segs[s].slot_base, segs[s].stride,
                       d_steps.as<uint32_t>() + segs[s].first,
                       d_special.as<uint8_t>() + segs[s].first, segs[s].count, 0,
                       seg_use64[s], slot_shift};
  }

  void plan(dpah_context* ctx, const dpah_job* jobs, const std::vector<uint64_t>& members) {
    const dpah_job& j0 = jobs[members[0]];
    record = j0.params.record != 0;
    for (uint64_t m : members) {
      const dpah_job& j = jobs[m];
      JobPlan p;
      p.job = m;
      p.base = N;
      p.count = j.count;
      p.use64 = j.params.use_64_bit_indices ? 1 : 0;
      // the caller's arrays are read here and by the uploads below only
      const uint64_t* rin = j.remaining;
      if (rin) {
        for (uint64_t i = 0; i < j.count; ++i) {
          DPAH_REQUIRE(rin[i] <= (1ull << 32) - 3, "remaining iterations must be below 2^32 - 2");
          p.rem_sum += rin[i];
        }
      } else if (j.count) {
        DPAH_REQUIRE(j.uniform_remaining <= (1ull << 32) - 3,
                     "remaining iterations must be below 2^32 - 2");
        p.rem_sum = j.uniform_remaining * j.count;
      }
      auto max_rem = [&](uint64_t first, uint64_t count) -> uint64_t {
        if (!rin) return count ? j.uniform_remaining : 0;
        uint64_t mx = 0;
        for (uint64_t i = 0; i < count; ++i) mx = std::max(mx, rin[first + i]);
        return mx;
      };
      if (record) {
        uint64_t active = j.count;
        do {   // pipeline.cpp:100-146 round order, tail first
          RoundPlan r;
          r.count = std::min(j.params.particles_per_round, active);
          r.first = active - r.count;
          r.stride = max_rem(r.first, r.count) + 2;
          p.rounds.push_back(r);
          active = r.first;
        } while (active > 0);
      } else if (j.count) {
        RoundPlan r;
        r.count = j.count;
        r.stride = max_rem(0, j.count);
        p.rounds.push_back(r);   // ordering key only; no slots
      }
      N += j.count;
      plans.push_back(std::move(p));
    }

    // segment table, longest rows first
    struct Ref { uint64_t plan, round; };
    std::vector<Ref> refs;
    for (uint64_t pi = 0; pi < plans.size(); ++pi)
      for (uint64_t ri = 0; ri < plans[pi].rounds.size(); ++ri)
        if (plans[pi].rounds[ri].count) refs.push_back({pi, ri});
    std::stable_sort(refs.begin(), refs.end(), [&](const Ref& a, const Ref& b) {
      return plans[a.plan].rounds[a.round].stride > plans[b.plan].rounds[b.round].stride;
    });
    segs.resize(refs.size());
    seg_use64.assign(refs.size(), 0);
    for (size_t s = 0; s < refs.size(); ++s) {
      RoundPlan& r = plans[refs[s].plan].rounds[refs[s].round];
      r.seg = (int)s;
      seg_use64[s] = plans[refs[s].plan].use64;
      segs[s].qbase = q;
      segs[s].first = plans[refs[s].plan].base + r.first;
      segs[s].count = r.count;
      segs[s].stride = record ? r.stride : 0;
      q += tiled_rows(r.count);   // segments start on 64-ordinal boundaries (whole waves)
    }
    // one slot layout for the launch; a batch's recorded launch stays
    // persistent unless the tuning asks for one thread per row
    const bool persistent =
        !record || !record_grid(ctx, j0.field->view, j0.params.integrator, true);
    slot_shift = record ? record_slot_shift(ctx, j0.field->view, j0.params.integrator, q,
                                            persistent)
                        : 0u;
    for (size_t s = 0; s < refs.size(); ++s) {
      segs[s].slot_base = slot_floats;
      if (record) slot_floats += slot_region_floats(segs[s].count, segs[s].stride, slot_shift);
    }
    row_off.assign(segs.size(), 0);
    sums_off.assign(segs.size(), 0);
    if (record) {
      for (size_t s = 0; s < segs.size(); ++s) {
        row_off[s] = rows_total;
        sums_off[s] = sums_total;
        rows_total += segs[s].count;
        sums_total += compact_sums_words(segs[s].count);
      }
    }

    const uint64_t n1 = std::max<uint64_t>(N, 1);
    d_seeds = DeviceBuffer(ctx, 3 * n1 * sizeof(float));
    d_rem = DeviceBuffer(ctx, n1 * sizeof(uint64_t));
    d_pos = DeviceBuffer(ctx, 3 * n1 * sizeof(float));
    d_rem_out = DeviceBuffer(ctx, n1 * sizeof(uint64_t));
    d_status = DeviceBuffer(ctx, n1);
    d_steps = DeviceBuffer(ctx, n1 * sizeof(uint32_t));
    d_special = DeviceBuffer(ctx, n1);
    d_segs = DeviceBuffer(ctx, std::max<size_t>(segs.size(), 1) * sizeof(Segment));
    if (record) d_slots = DeviceBuffer(ctx, std::max<uint64_t>(slot_floats, 1) * sizeof(float));
    if (record && rows_total) {
      d_vcount = DeviceBuffer(ctx, rows_total * sizeof(uint64_t));
      d_icount = DeviceBuffer(ctx, rows_total * sizeof(uint64_t));
      d_sums = DeviceBuffer(ctx, sums_total * sizeof(uint64_t));
    }
    // each job's seeds and iterations straight from the caller's arrays
    for (const JobPlan& p : plans) {
      const dpah_job& j = jobs[p.job];
      if (!j.count) continue;
      DPAH_HIP(hipMemcpyAsync(d_seeds.as<float>() + 3 * p.base, j.seeds,
                              3 * j.count * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      if (j.remaining) {
        DPAH_HIP(hipMemcpyAsync(d_rem.as<uint64_t>() + p.base, j.remaining,
                                j.count * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
      } else {
        const unsigned grid = (unsigned)std::min<uint64_t>((j.count + 255) / 256, 1024);
        hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, ctx->stream,
                           d_rem.as<uint64_t>() + p.base, j.count, j.uniform_remaining);
        DPAH_HIP(hipGetLastError());
      }
    }
    if (!segs.empty())
      DPAH_HIP(hipMemcpyAsync(d_segs.ptr, segs.data(), segs.size() * sizeof(Segment),
                              hipMemcpyHostToDevice, ctx->stream));

    L.f = &j0.field->view;
    L.seed_kind = kSeedArray;
    L.segs = d_segs.as<Segment>();
    L.nseg = (uint32_t)segs.size();
    L.seeds = d_seeds.as<float>();
    L.rem_in = d_rem.as<uint64_t>();
    L.pos_out = d_pos.as<float>();
    L.rem_out = d_rem_out.as<uint64_t>();
    L.status_out = d_status.as<uint8_t>();
    L.steps_out = record ? d_steps.as<uint32_t>() : nullptr;
    L.special_out = record ? d_special.as<uint8_t>() : nullptr;
    L.slots = record ? d_slots.as<float>() : nullptr;
    L.slot_shift = slot_shift;
    L.n = q;
    L.dt = j0.params.step_size;
    L.t0 = j0.params.start_time;
    L.integrator = j0.params.integrator;
    L.persistent = persistent;
  }

  void enqueue(dpah_context* ctx) const {
    if (L.nseg) launch_advect(ctx, L);
    if (record && rows_total)
      for (size_t s = 0; s < segs.size(); ++s)
        compact_count(ctx, rows_of(s), d_vcount.as<uint64_t>() + row_off[s],
                      d_icount.as<uint64_t>() + row_off[s], d_sums.as<uint64_t>() + sums_off[s],
                      ctx->stream);
  }

  void finish(dpah_context* ctx, std::vector<std::unique_ptr<dpah_result>>& out) const {
    // totals and final states in one read-back: [sums | remaining | positions
    // | status], in the context's pinned staging when it fits
    const uint64_t bytes = 8 * sums_total + 21 * N;
    std::unique_ptr<uint8_t[]> heap;
    uint8_t* stage = static_cast<uint8_t*>(ctx->host_staging(std::max<uint64_t>(bytes, 1)));
    if (!stage) {
      heap.reset(new uint8_t[bytes]);
      stage = heap.get();
    }
    const uint64_t* h_sums = reinterpret_cast<const uint64_t*>(stage);
    const uint64_t* h_rem = reinterpret_cast<const uint64_t*>(stage + 8 * sums_total);
    const float* h_pos = reinterpret_cast<const float*>(stage + 8 * sums_total + 8 * N);
    const uint8_t* h_status = stage + 8 * sums_total + 20 * N;
    if (record && rows_total)
      DPAH_HIP(hipMemcpyAsync(stage, d_sums.ptr, sums_total * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
    if (N) {
      DPAH_HIP(hipMemcpyAsync(stage + 8 * sums_total, d_rem_out.ptr, N * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
      DPAH_HIP(hipMemcpyAsync(stage + 8 * sums_total + 8 * N, d_pos.ptr, 3 * N * sizeof(float),
                              hipMemcpyDeviceToHost, ctx->stream));
      DPAH_HIP(hipMemcpyAsync(stage + 8 * sums_total + 20 * N, d_status.ptr, N,
                              hipMemcpyDeviceToHost, ctx->stream));
    }
    ctx->sync();

    // every round's writes first, in one launch, so the device works while
    // the host copies the final states out of the staging below
    std::vector<WriteJob> writes;
    if (record) {
      for (const JobPlan& p : plans)
        for (const RoundPlan& r : p.rounds) {
          if (!r.count) continue;
          const size_t s = (size_t)r.seg;
          const uint64_t* tot = h_sums + sums_off[s] + compact_sums_words(segs[s].count) - 2;
          if (!p.use64)
            DPAH_REQUIRE(tot[0] <= 0xffffffffull,
                         "round has more than 2^32-1 vertices; 32-bit indices cannot address it");
          writes.push_back(WriteJob{rows_of(s), d_vcount.as<uint64_t>() + row_off[s],
                                    d_icount.as<uint64_t>() + row_off[s], tot[0], tot[1]});
        }
    }
    DeviceBuffer table;
    std::vector<Chunk> chunks = compact_write_many(ctx, writes, table, ctx->stream);
    size_t w = 0;
    for (const JobPlan& p : plans) {
      auto res = std::make_unique<dpah_result>();
      res->ctx = ctx;
      res->use64 = p.use64;
      if (record) {
        for (const RoundPlan& r : p.rounds) {
          Round round;
          if (r.count) {
            const Chunk& ch = chunks[w++];
            round.V = ch.V;
            round.I = ch.I;
            round.chunks.push_back(ch);
          }
          res->rounds.push_back(std::move(round));
        }
      }
      out[p.job] = std::move(res);
    }
    for (const JobPlan& p : plans) {
      dpah_result* res = out[p.job].get();
      res->final_pos.assign(h_pos + 3 * p.base, h_pos + 3 * (p.base + p.count));
      res->final_rem.assign(h_rem + p.base, h_rem + p.base + p.count);
      res->final_status.assign(h_status + p.base, h_status + p.base + p.count);
      uint64_t left = 0;
      for (uint64_t i = 0; i < p.count; ++i) left += res->final_rem[i];
      res->total_steps = p.rem_sum - left;
    }
    if (record) ctx->sync();   // the chunks are complete when the results are handed out
  }
};

// Groups jobs by what the kernel takes as launch-uniform arguments.
std::vector<std::vector<uint64_t>> group_jobs(const dpah_job* jobs, uint64_t job_count) {
  std::map<std::tuple<const dpah_field*, uint32_t, uint32_t, int, int>, std::vector<uint64_t>>
      groups;
  for (uint64_t i = 0; i < job_count; ++i) {
    const dpah_job& j = jobs[i];
    groups[{j.field, float_bits(j.params.step_size), float_bits(j.params.start_time),
            j.params.integrator, j.params.record ? 1 : 0}]
        .push_back(i);
  }
  std::vector<std::vector<uint64_t>> out;
  for (auto& g : groups) out.push_back(std::move(g.second));
  return out;
}

}  // namespace

DPAH_EXPORT int dpah_advect_batch(dpah_context* ctx, const dpah_job* jobs, uint64_t job_count,
                                  dpah_result** results) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && (job_count == 0 || (jobs && results)), "null argument");
    for (uint64_t i = 0; i < job_count; ++i) validate(ctx, jobs[i]);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    std::vector<std::unique_ptr<dpah_result>> out(job_count);
    for (const auto& members : group_jobs(jobs, job_count)) {
      GroupRun g;
      g.plan(ctx, jobs, members);
      g.enqueue(ctx);
      g.finish(ctx, out);
    }
    for (uint64_t i = 0; i < job_count; ++i) results[i] = out[i].release();
  });
}

// A batch planned and uploaded once and captured into one hipGraph (every
// group's counter reset, persistent launch and round counts / scans, in group
// order on the context's stream).  Replays recompute every output from the
// uploaded seeds; dpah_batch_pass_results then does the part whose sizes the
// replay decides (read-back of totals and final states, the round writes).
struct dpah_batch_pass {
  dpah_context* ctx = nullptr;
  uint64_t job_count = 0;
  std::vector<std::unique_ptr<GroupRun>> groups;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t launches = 0;
  ~dpah_batch_pass() {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);   // uploads / replays done before frees
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

DPAH_EXPORT int dpah_batch_pass_create(dpah_context* ctx, const dpah_job* jobs,
                                       uint64_t job_count, dpah_batch_pass** out) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && jobs && out, "null argument");
    DPAH_REQUIRE(job_count > 0, "a batch pass needs at least one job");
    for (uint64_t i = 0; i < job_count; ++i) validate(ctx, jobs[i]);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    auto p = std::make_unique<dpah_batch_pass>();
    p->ctx = ctx;
    p->job_count = job_count;
    for (const auto& members : group_jobs(jobs, job_count)) {
      p->groups.push_back(std::make_unique<GroupRun>());
      p->groups.back()->plan(ctx, jobs, members);
    }
    ctx->sync();   // uploads done; nothing of the caller's in flight while capturing
    const bool timing = ctx->timing;
    ctx->timing = false;   // no event records inside the graph
    try {
      DPAH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      try {
        for (const auto& g : p->groups) g->enqueue(ctx);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      DPAH_HIP(hipStreamEndCapture(ctx->stream, &p->graph));
    } catch (...) {
      ctx->timing = timing;
      throw;
    }
    ctx->timing = timing;
    DPAH_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    *out = p.release();
  });
}

DPAH_EXPORT int dpah_batch_pass_launch(dpah_batch_pass* p) {
  return guarded([&] {
    DPAH_REQUIRE(p, "null argument");
    dpah_context* ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ctx->timed("advect_batch_graph", [&] { DPAH_HIP(hipGraphLaunch(p->exec, ctx->stream)); });
    p->launches += 1;
  });
}

DPAH_EXPORT int dpah_batch_pass_results(dpah_batch_pass* p, dpah_result** results) {
  return guarded([&] {
    DPAH_REQUIRE(p && results, "null argument");
    DPAH_REQUIRE(p->launches > 0, "batch pass has not been launched");
    dpah_context* ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    std::vector<std::unique_ptr<dpah_result>> out(p->job_count);
    for (const auto& g : p->groups) g->finish(ctx, out);
    for (uint64_t i = 0; i < p->job_count; ++i) results[i] = out[i].release();
  });
}

DPAH_EXPORT int dpah_batch_pass_destroy(dpah_batch_pass* p) {
  return guarded([&] {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> lk(p->ctx->mu);
    (void)hipSetDevice(p->ctx->device);
    delete p;
  });
}

// Launch shapes of a batch's recorded launches: per group of jobs, workgroups
// per CU x refill x slot layout (one-slot runs, or 8-slot blocks staged in
// LDS), each timed on the group's own persistent launch (the slot layout is
// planned again for each candidate, since it decides the slot buffer).
DPAH_EXPORT int dpah_autotune_batch(dpah_context* ctx, const dpah_job* jobs, uint64_t job_count,
                                    dpah_tuning* best_out, double* best_ms) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && jobs && job_count > 0, "null argument or empty batch");
    for (uint64_t i = 0; i < job_count; ++i) validate(ctx, jobs[i]);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    uint64_t most = 0;
    for (const auto& members : group_jobs(jobs, job_count)) {
      const dpah_job& j0 = jobs[members[0]];
      const std::string key = launch_key(j0.field->view, j0.params.integrator, kSeedArray,
                                         j0.params.record != 0);
      uint64_t work = 0;
      for (uint64_t m : members) work += jobs[m].count;
      std::vector<Tuning> cands;
      for (int bpc : {3, 4, 5, 6, 8})
        for (int refill : {16, 64})
          for (int stage : {1, 2}) {
            if (!j0.params.record && stage == 2) continue;
            Tuning t = ctx->tune;
            t.block = 256;
            t.blocks_per_cu = bpc;
            t.refill = refill;
            t.slot_stage = j0.params.record ? stage : 0;
            cands.push_back(t);
          }
      if (j0.params.record) {   // and one thread per row (no launch shape to vary)
        Tuning t = ctx->tune;
        t.block = 256;
        t.slot_stage = 3;
        cands.push_back(t);
      }
      std::unique_ptr<GroupRun> g;
      auto prepare = [&](const Tuning&) {
        g.reset();
        ctx->sync();   // the previous candidate's buffers are idle before reuse
        g = std::make_unique<GroupRun>();
        g->plan(ctx, jobs, members);
      };
      double ms = 0;
      const Tuning best = tune_key(ctx, key, cands, prepare, [&] {
        if (g->L.nseg) launch_advect(ctx, g->L);
      }, &ms);
      g.reset();
      ctx->sync();
      if (work >= most) {   // report the group with the most particles
        most = work;
        if (best_out) tuning_to_abi(best, best_out);
        if (best_ms) *best_ms = ms;
      }
    }
  });
}
