// tune.hip -- the launch-shape autotuner of the persistent advect kernel.
//
// dpah_autotune_endpoints times every candidate tuning (wave refill threshold,
// workgroups per CU, queue ordering, field layout) on a caller-supplied sample
// of an endpoint workload, keeps the fastest for that launch key (field shape,
// integrator, seed kind, record flag) and, when the context has a cache file,
// writes it there so later contexts start tuned.  Results never depend on the
// tuning: every candidate runs the same per-particle arithmetic.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>

#include "runtime.hpp"

using namespace dpah;

namespace {

struct Candidate {
  int block, refill, blocks_per_cu, order, cells, order_brick;
};

// Cache file: one line per launch key,
// "<arch> <cus> <key> <block> <blocks_per_cu> <order> <parts> <refill> <cells> <fastdiv>
// <cell_brick> <order_brick> <ms>" (cell_brick is a field-creation setting, kept for the
// record).
void save_cache(const dpah_context* ctx, const std::string& key, const Tuning& t, double ms) {
  if (ctx->autotune_path.empty()) return;
  std::vector<std::string> keep;
  {
    std::ifstream in(ctx->autotune_path);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ss(line);
      std::string arch, k;
      int cus = 0;
      if (ss >> arch >> cus >> k && !(arch == ctx->arch && cus == ctx->cu_count && k == key))
        keep.push_back(line);
    }
  }
  std::ofstream out(ctx->autotune_path, std::ios::trunc);
  for (const auto& l : keep) out << l << "\n";
  out << ctx->arch << " " << ctx->cu_count << " " << key << " " << t.block << " "
      << t.blocks_per_cu << " " << t.order << " " << t.parts << " " << t.refill << " " << t.cells
      << " " << t.fastdiv << " " << t.cell_brick << " " << t.order_brick << " " << ms << "\n";
  if (!out) throw Error(DPAH_ERR_INVALID, "cannot write the autotune cache " + ctx->autotune_path);
}

}  // namespace

namespace dpah {

// Entries of the cache file for this device go into ctx->tuned.
void load_autotune_cache(dpah_context* ctx) {
  if (ctx->autotune_path.empty()) return;
  std::ifstream in(ctx->autotune_path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string arch, key;
    int cus = 0;
    Tuning t;
    double ms = 0;
    if (!(ss >> arch >> cus >> key >> t.block >> t.blocks_per_cu >> t.order >> t.parts >>
          t.refill >> t.cells >> t.fastdiv >> t.cell_brick >> t.order_brick >> ms))
      continue;
    if (arch != ctx->arch || cus != ctx->cu_count) continue;
    const bool ok = t.block >= 64 && t.block <= 256 && t.block % 64 == 0 &&
                    t.blocks_per_cu >= 0 && t.blocks_per_cu <= 32 && (t.order == 0 || t.order == 1) &&
                    t.parts >= 1 && t.parts <= kMaxParts && t.refill >= 0 && t.refill <= 64 &&
                    t.cells >= 0 && t.cells <= 2 && (t.fastdiv == 0 || t.fastdiv == 1) &&
                    (t.cell_brick == 0 || (t.cell_brick >= 2 && t.cell_brick <= 5)) &&
                    t.order_brick >= 0 && t.order_brick <= 5;
    if (ok) ctx->tuned[key] = t;
  }
}

}  // namespace dpah

DPAH_EXPORT int dpah_autotune_endpoints(dpah_context* ctx, const dpah_field* field,
                                        const float* seeds, uint64_t count,
                                        uint64_t uniform_remaining, float step_size,
                                        float start_time, int integrator, dpah_tuning* best_out,
                                        double* best_ms) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field && seeds && count > 0, "null argument or empty sample");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
                 "unknown integrator");
    DPAH_REQUIRE(uniform_remaining <= (1ull << 32) - 3,
                 "remaining iterations must be below 2^32 - 2");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    const std::string key = launch_key(field->view, integrator, kSeedArray, false);
    DeviceBuffer pos(ctx, count * 3 * sizeof(float));
    AdvectLaunch L{};
    L.f = &field->view;
    L.seeds = seeds;
    L.uniform_rem = uniform_remaining;
    L.pos_out = pos.as<float>();
    L.n = count;
    L.dt = step_size;
    L.t0 = start_time;
    L.integrator = integrator;
    L.persistent = true;
    hipEvent_t e0 = ctx->take_event(), e1 = ctx->take_event();
    auto time_once = [&](const Tuning& t) {
      ctx->tuned[key] = t;
      DPAH_HIP(hipEventRecord(e0, ctx->stream));
      launch_advect(ctx, L);
      DPAH_HIP(hipEventRecord(e1, ctx->stream));
      DPAH_HIP(hipEventSynchronize(e1));
      float ms = 0;
      DPAH_HIP(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms;
    };
    const Tuning base = ctx->tune;
    std::vector<Candidate> cands;
    // workgroup size x workgroups per CU as waves per CU (12..32), refill,
    // queue order with its brick size, field layout
    for (int block : {128, 256})
      for (int waves : {12, 16, 20, 24, 32})
        for (int refill : {4, 8, 16, 64})
          for (int order : {0, 1})
            for (int obrick : {2, 3})
              for (int cells : {0, 1}) {
                if (cells && !field->view.cells) continue;
                if (!order && obrick != 2) continue;   // the brick only shapes the sort
                cands.push_back({block, refill, waves * 64 / block, order, cells, obrick});
              }
    Tuning best = base;
    double best_t = 1e300;
    try {
      time_once(base);   // warm-up: code objects, caches, allocator blocks
      for (const Candidate& c : cands) {
        Tuning t = base;
        t.block = c.block;
        t.order_brick = c.order_brick;
        t.refill = c.refill;
        t.blocks_per_cu = c.blocks_per_cu;
        t.order = c.order;
        t.cells = c.cells;
        const double ms = std::min(time_once(t), time_once(t));
        if (ms < best_t) { best_t = ms; best = t; }
      }
    } catch (...) {
      ctx->tuned.erase(key);
      ctx->event_pool.push_back(e0);
      ctx->event_pool.push_back(e1);
      throw;
    }
    ctx->event_pool.push_back(e0);
    ctx->event_pool.push_back(e1);
    ctx->tuned[key] = best;
    save_cache(ctx, key, best, best_t);
    if (best_out) {
      best_out->block = best.block;
      best_out->blocks_per_cu = best.blocks_per_cu;
      best_out->order = best.order;
      best_out->parts = best.parts;
      best_out->refill = best.refill;
      best_out->cells = best.cells;
      best_out->fastdiv = best.fastdiv;
      best_out->cell_brick = best.cell_brick;
      best_out->order_brick = best.order_brick;
    }
    if (best_ms) *best_ms = best_t;
  });
}
