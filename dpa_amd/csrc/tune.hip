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
#include <functional>
#include <sstream>

#include "runtime.hpp"

using namespace dpah;

namespace {

// Cache file: one line per launch key,
// "<arch> <cus> <key> <block> <blocks_per_cu> <order> <parts> <refill> <cells> <fastdiv>
// <cell_brick> <order_brick> <ms> [<lattice_brick> <slot_stage>]" (cell_brick is a
// field-creation setting, kept for the record; lines without the last two fields
// read them as 0).
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
      << " " << t.fastdiv << " " << t.cell_brick << " " << t.order_brick << " " << ms << " "
      << t.lattice_brick << " " << t.slot_stage << " " << t.lattice_group << " " << t.node_brick
      << "\n";
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
    int lb = 0, st = 0, lg = 0;
    if (ss >> lb >> st) { t.lattice_brick = lb; t.slot_stage = st; }
    if (ss >> lg) t.lattice_group = lg;
    int nbk = 1;
    if (ss >> nbk) t.node_brick = nbk;
    if (arch != ctx->arch || cus != ctx->cu_count) continue;
    const bool ok = t.block >= 64 && t.block <= 256 && t.block % 64 == 0 &&
                    t.blocks_per_cu >= 0 && t.blocks_per_cu <= 32 && (t.order == 0 || t.order == 1) &&
                    t.parts >= 1 && t.parts <= kMaxParts && t.refill >= 0 && t.refill <= 64 &&
                    t.cells >= 0 && t.cells <= 2 && (t.fastdiv == 0 || t.fastdiv == 1) &&
                    (t.cell_brick == 0 || (t.cell_brick >= 2 && t.cell_brick <= 5)) &&
                    t.order_brick >= 0 && t.order_brick <= 5 &&
                    (t.lattice_brick == 0 || t.lattice_brick == 2 || t.lattice_brick == 4 ||
                     t.lattice_brick == 8) &&
                    t.slot_stage >= 0 && t.slot_stage <= 3 &&
                    (t.lattice_group == 0 || t.lattice_group == 1 || t.lattice_group == 2 ||
                     t.lattice_group == 4 || t.lattice_group == 8 || t.lattice_group == 16) &&
                    t.node_brick >= 0 && t.node_brick <= 2;
    if (ok) ctx->tuned[key] = t;
  }
}

void tuning_to_abi(const Tuning& t, dpah_tuning* out) {
  out->block = t.block;
  out->blocks_per_cu = t.blocks_per_cu;
  out->order = t.order;
  out->parts = t.parts;
  out->refill = t.refill;
  out->cells = t.cells;
  out->fastdiv = t.fastdiv;
  out->cell_brick = t.cell_brick;
  out->order_brick = t.order_brick;
  out->lattice_brick = t.lattice_brick;
  out->slot_stage = t.slot_stage;
  out->lattice_group = t.lattice_group;
  out->node_brick = t.node_brick;
}

// Times `launch` (one launch on the context's stream) under every candidate
// tuning of `key`, twice each after a warm-up, keeps the fastest in ctx->tuned
// and the cache file, and returns it.  On failure the key's entry is dropped.
Tuning tune_key(dpah_context* ctx, const std::string& key, const std::vector<Tuning>& cands,
                const std::function<void(const Tuning&)>& prepare,
                const std::function<void()>& launch, double* best_ms) {
  hipEvent_t e0 = ctx->take_event(), e1 = ctx->take_event();
  auto time_once = [&]() {
    DPAH_HIP(hipEventRecord(e0, ctx->stream));
    launch();
    DPAH_HIP(hipEventRecord(e1, ctx->stream));
    DPAH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    DPAH_HIP(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms;
  };
  Tuning best = ctx->tune;
  double best_t = 1e300;
  try {
    ctx->tuned[key] = ctx->tune;
    prepare(ctx->tune);
    time_once();   // warm-up: code objects, caches, allocator blocks
    for (const Tuning& t : cands) {
      ctx->tuned[key] = t;
      prepare(t);
      const double ms = std::min(time_once(), time_once());
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
  if (best_ms) *best_ms = best_t;
  return best;
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
    // workgroup size x workgroups per CU as waves per CU (12..32), refill,
    // queue order with its brick size, field layout
    std::vector<Tuning> cands;
    for (int block : {128, 256})
      for (int waves : {12, 16, 20, 24, 32})
        for (int refill : {4, 8, 16, 64})
          for (int order : {0, 1})
            for (int obrick : {2, 3})
              for (int cells : {0, 1}) {
                if (cells && !field->view.cells) continue;
                if (!order && obrick != 2) continue;   // the brick only shapes the sort
                Tuning t = ctx->tune;
                t.block = block;
                t.blocks_per_cu = waves * 64 / block;
                t.refill = refill;
                t.order = order;
                t.order_brick = obrick;
                t.cells = cells;
                cands.push_back(t);
              }
    const Tuning best = tune_key(ctx, key, cands, [](const Tuning&) {},
                                 [&] { launch_advect(ctx, L); }, best_ms);
    if (best_out) tuning_to_abi(best, best_out);
  });
}

DPAH_EXPORT int dpah_autotune_lattice(dpah_context* ctx, const dpah_field* field,
                                      const float origin[3], const float spacing[3],
                                      const uint64_t shape[3], const float* delta,
                                      uint64_t iterations, float step_size, float start_time,
                                      int integrator, dpah_tuning* best_out, double* best_ms) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field && origin && spacing && shape, "null argument");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
                 "unknown integrator");
    DPAH_REQUIRE(iterations <= (1ull << 32) - 3, "iterations must be below 2^32 - 2");
    const uint64_t points = shape[0] * shape[1] * shape[2];
    DPAH_REQUIRE(points > 0, "empty lattice");
    SeedLattice lat{};
    uint64_t n = points;
    int kind = kSeedLattice;
    if (delta) {
      kind = aux_lattice(lat, origin, spacing, shape, delta, &n);
    } else {
      lat.lo0 = origin[0]; lat.lo1 = origin[1]; lat.lo2 = origin[2];
      lat.st0 = spacing[0]; lat.st1 = spacing[1]; lat.st2 = spacing[2];
      lat.d0 = shape[0]; lat.d1 = shape[1]; lat.d2 = shape[2];
    }
    DPAH_REQUIRE(n < (1ull << 32), "lattice too large to tune in one launch");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    const std::string key = launch_key(field->view, integrator, kind, false);
    DeviceBuffer pos(ctx, n * 3 * sizeof(float));
    AdvectLaunch L{};
    L.f = &field->view;
    L.seed_kind = kind;
    L.lat = lat;
    L.uniform_rem = iterations;
    L.pos_out = pos.as<float>();
    L.n = n;
    L.dt = step_size;
    L.t0 = start_time;
    L.integrator = integrator;
    L.persistent = true;
    // workgroups per CU x refill x queue walk (C order; bricks of 2 / 4 / 8
    // points per axis, alone or in groups of 2 or 4 bricks per axis)
    const int walks[][2] = {{0, 1}, {2, 1}, {4, 1}, {8, 1}, {2, 2}, {4, 2}, {4, 4}};
    std::vector<Tuning> cands;
    for (int bpc : {2, 3, 4, 6})
      for (int refill : {32, 64})
        for (const auto& w : walks) {
          Tuning t = ctx->tune;
          t.block = 256;
          t.blocks_per_cu = bpc;
          t.refill = refill;
          t.order = w[0] ? 1 : 0;
          t.lattice_brick = w[0];
          t.lattice_group = w[1];
          cands.push_back(t);
        }
    const Tuning best = tune_key(ctx, key, cands, [](const Tuning&) {},
                                 [&] { launch_advect(ctx, L); }, best_ms);
    if (best_out) tuning_to_abi(best, best_out);
  });
}
