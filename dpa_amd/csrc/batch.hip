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
  std::vector<uint64_t> rem;
  std::vector<RoundPlan> rounds;
};

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

// Runs one group of jobs that share field, step size, start time, integrator
// and record flag: one persistent launch, then per-round compaction.
void run_group(dpah_context* ctx, const dpah_job* jobs, const std::vector<uint64_t>& members,
               std::vector<std::unique_ptr<dpah_result>>& out) {
  const dpah_job& j0 = jobs[members[0]];
  const bool record = j0.params.record != 0;
  std::vector<JobPlan> plans;
  uint64_t N = 0;
  for (uint64_t m : members) {
    const dpah_job& j = jobs[m];
    JobPlan p;
    p.job = m;
    p.base = N;
    p.rem.resize(j.count);
    for (uint64_t i = 0; i < j.count; ++i) {
      p.rem[i] = j.remaining ? j.remaining[i] : j.uniform_remaining;
      DPAH_REQUIRE(p.rem[i] <= (1ull << 32) - 3, "remaining iterations must be below 2^32 - 2");
    }
    if (record) {
      uint64_t active = j.count;
      do {   // pipeline.cpp:100-146 round order, tail first
        RoundPlan r;
        r.count = std::min(j.params.particles_per_round, active);
        r.first = active - r.count;
        uint64_t mx = 0;
        for (uint64_t i = 0; i < r.count; ++i) mx = std::max(mx, p.rem[r.first + i]);
        r.stride = mx + 2;
        p.rounds.push_back(r);
        active = r.first;
      } while (active > 0);
    } else if (j.count) {
      RoundPlan r;
      r.count = j.count;
      for (uint64_t i = 0; i < j.count; ++i) r.stride = std::max(r.stride, p.rem[i]);
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
  std::vector<Segment> segs(refs.size());
  uint64_t q = 0, slot_floats = 0;
  for (size_t s = 0; s < refs.size(); ++s) {
    RoundPlan& r = plans[refs[s].plan].rounds[refs[s].round];
    r.seg = (int)s;
    segs[s].qbase = q;
    segs[s].first = plans[refs[s].plan].base + r.first;
    segs[s].count = r.count;
    segs[s].stride = record ? r.stride : 0;
    segs[s].slot_base = slot_floats;
    q += tiled_rows(r.count);   // segments start on 64-ordinal boundaries (whole waves)
    if (record) slot_floats += 3 * tiled_rows(r.count) * r.stride;
  }

  const uint64_t n1 = std::max<uint64_t>(N, 1);
  std::vector<float> h_seeds(3 * n1);
  std::vector<uint64_t> h_rem(n1);
  for (const JobPlan& p : plans) {
    const dpah_job& j = jobs[p.job];
    if (j.count) std::memcpy(h_seeds.data() + 3 * p.base, j.seeds, 3 * j.count * sizeof(float));
    std::copy(p.rem.begin(), p.rem.end(), h_rem.begin() + p.base);
  }
  DeviceBuffer d_seeds(ctx, 3 * n1 * sizeof(float));
  DeviceBuffer d_rem(ctx, n1 * sizeof(uint64_t));
  DeviceBuffer d_pos(ctx, 3 * n1 * sizeof(float));
  DeviceBuffer d_rem_out(ctx, n1 * sizeof(uint64_t));
  DeviceBuffer d_status(ctx, n1);
  DeviceBuffer d_steps(ctx, n1 * sizeof(uint32_t));
  DeviceBuffer d_special(ctx, n1);
  DeviceBuffer d_segs(ctx, std::max<size_t>(segs.size(), 1) * sizeof(Segment));
  DeviceBuffer d_slots;
  if (record) d_slots = DeviceBuffer(ctx, std::max<uint64_t>(slot_floats, 1) * sizeof(float));
  DPAH_HIP(hipMemcpyAsync(d_seeds.ptr, h_seeds.data(), 3 * N * sizeof(float),
                          hipMemcpyHostToDevice, ctx->stream));
  DPAH_HIP(hipMemcpyAsync(d_rem.ptr, h_rem.data(), N * sizeof(uint64_t), hipMemcpyHostToDevice,
                          ctx->stream));
  if (!segs.empty())
    DPAH_HIP(hipMemcpyAsync(d_segs.ptr, segs.data(), segs.size() * sizeof(Segment),
                            hipMemcpyHostToDevice, ctx->stream));

  AdvectLaunch L{};
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
  L.n = q;
  L.dt = j0.params.step_size;
  L.t0 = j0.params.start_time;
  L.integrator = j0.params.integrator;
  L.persistent = true;
  if (L.nseg) launch_advect(ctx, L);

  // Every recorded segment's counts and scan go on the stream behind the
  // launch; their totals come back with the final states in one
  // synchronisation, then every write is launched and the group waits once.
  std::vector<uint64_t> row_off(segs.size(), 0), sums_off(segs.size(), 0);
  uint64_t rows_total = 0, sums_total = 0;
  if (record) {
    for (size_t s = 0; s < segs.size(); ++s) {
      row_off[s] = rows_total;
      sums_off[s] = sums_total;
      rows_total += segs[s].count;
      sums_total += compact_sums_words(segs[s].count);
    }
  }
  DeviceBuffer d_vcount, d_icount, d_sums;
  std::vector<uint64_t> h_sums(sums_total);
  auto rows_of = [&](size_t s) {
    return CompactRows{d_slots.as<float>() + segs[s].slot_base, segs[s].stride,
                       d_steps.as<uint32_t>() + segs[s].first,
                       d_special.as<uint8_t>() + segs[s].first, segs[s].count, 0, 0};
  };
  std::vector<int> seg_use64(segs.size(), 0);
  for (const JobPlan& p : plans)
    for (const RoundPlan& r : p.rounds)
      if (r.seg >= 0) seg_use64[r.seg] = jobs[p.job].params.use_64_bit_indices ? 1 : 0;
  if (record && rows_total) {
    d_vcount = DeviceBuffer(ctx, rows_total * sizeof(uint64_t));
    d_icount = DeviceBuffer(ctx, rows_total * sizeof(uint64_t));
    d_sums = DeviceBuffer(ctx, sums_total * sizeof(uint64_t));
    for (size_t s = 0; s < segs.size(); ++s) {
      CompactRows r = rows_of(s);
      r.use64 = seg_use64[s];
      compact_count(ctx, r, d_vcount.as<uint64_t>() + row_off[s],
                    d_icount.as<uint64_t>() + row_off[s], d_sums.as<uint64_t>() + sums_off[s],
                    ctx->stream);
    }
    DPAH_HIP(hipMemcpyAsync(h_sums.data(), d_sums.ptr, sums_total * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, ctx->stream));
  }

  std::vector<float> h_pos(3 * n1);
  std::vector<uint64_t> h_rem_out(n1);
  std::vector<uint8_t> h_status(n1);
  DPAH_HIP(hipMemcpyAsync(h_pos.data(), d_pos.ptr, 3 * N * sizeof(float), hipMemcpyDeviceToHost,
                          ctx->stream));
  DPAH_HIP(hipMemcpyAsync(h_rem_out.data(), d_rem_out.ptr, N * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  DPAH_HIP(hipMemcpyAsync(h_status.data(), d_status.ptr, N, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();

  for (const JobPlan& p : plans) {
    const dpah_job& j = jobs[p.job];
    auto res = std::make_unique<dpah_result>();
    res->ctx = ctx;
    res->use64 = j.params.use_64_bit_indices ? 1 : 0;
    res->final_pos.assign(h_pos.begin() + 3 * p.base, h_pos.begin() + 3 * (p.base + j.count));
    res->final_rem.assign(h_rem_out.begin() + p.base, h_rem_out.begin() + p.base + j.count);
    res->final_status.assign(h_status.begin() + p.base, h_status.begin() + p.base + j.count);
    uint64_t steps = 0;
    for (uint64_t i = 0; i < j.count; ++i) steps += p.rem[i] - res->final_rem[i];
    res->total_steps = steps;
    if (record) {
      for (const RoundPlan& r : p.rounds) {
        Round round;
        if (r.count) {
          const size_t s = (size_t)r.seg;
          CompactRows rows = rows_of(s);
          rows.use64 = res->use64;
          const uint64_t* tot = h_sums.data() + sums_off[s] + compact_sums_words(segs[s].count) - 2;
          if (!res->use64)
            DPAH_REQUIRE(tot[0] <= 0xffffffffull,
                         "round has more than 2^32-1 vertices; 32-bit indices cannot address it");
          Chunk ch = compact_write(ctx, rows, d_vcount.as<uint64_t>() + row_off[s],
                                   d_icount.as<uint64_t>() + row_off[s], tot[0], tot[1],
                                   ctx->stream);
          round.V = ch.V;
          round.I = ch.I;
          round.chunks.push_back(ch);
        }
        res->rounds.push_back(std::move(round));
      }
    }
    out[p.job] = std::move(res);
  }
  if (record) ctx->sync();   // the chunks are complete when the results are handed out
}

}  // namespace

DPAH_EXPORT int dpah_advect_batch(dpah_context* ctx, const dpah_job* jobs, uint64_t job_count,
                                  dpah_result** results) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && (job_count == 0 || (jobs && results)), "null argument");
    for (uint64_t i = 0; i < job_count; ++i) validate(ctx, jobs[i]);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    // group by what the kernel takes as launch-uniform arguments
    std::map<std::tuple<const dpah_field*, uint32_t, uint32_t, int, int>, std::vector<uint64_t>>
        groups;
    for (uint64_t i = 0; i < job_count; ++i) {
      const dpah_job& j = jobs[i];
      groups[{j.field, float_bits(j.params.step_size), float_bits(j.params.start_time),
              j.params.integrator, j.params.record ? 1 : 0}]
          .push_back(i);
    }
    std::vector<std::unique_ptr<dpah_result>> out(job_count);
    for (auto& g : groups) run_group(ctx, jobs, g.second, out);
    for (uint64_t i = 0; i < job_count; ++i) results[i] = out[i].release();
  });
}
