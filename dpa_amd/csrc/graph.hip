// graph.hip -- endpoint passes captured once into a hipGraph and replayed.
//
// A pass is dpah_advect_endpoints_device with fixed arguments: the counter
// reset, the queue ordering (five kernels) and the persistent advect kernel.
// Capturing them once and replaying the instantiated graph removes the host
// launch work of every repetition (time stepping over the same seeds, FTLE
// over successive windows, benchmarks).  The pass owns the ordering scratch,
// so nothing is allocated while the graph is captured or replayed, and every
// buffer the graph touches lives as long as the pass.  The graph holds kernel
// nodes only: the counters are cleared by a kernel, not by a memset node.
#include "runtime.hpp"

using namespace dpah;

struct dpah_pass {
  dpah_context* ctx = nullptr;
  DeviceBuffer key, order, count;
  OrderScratch scratch;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t launches = 0;
};

static void destroy_pass(dpah_pass* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  delete p;
}

DPAH_EXPORT int dpah_pass_create_endpoints(dpah_context* ctx, const dpah_field* field,
                                           const float* seeds, const uint64_t* remaining,
                                           uint64_t uniform_remaining, uint64_t count,
                                           float step_size, float start_time, int integrator,
                                           float* positions_out, uint64_t* remaining_out,
                                           uint8_t* status_out, dpah_pass** out) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field && out, "null argument");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(count > 0 && seeds && positions_out, "seeds / positions_out is null or empty");
    DPAH_REQUIRE(remaining || uniform_remaining <= (1ull << 32) - 3,
                 "remaining iterations must be below 2^32 - 2");
    DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
                 "unknown integrator");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    auto* p = new dpah_pass();
    p->ctx = ctx;
    const bool timing = ctx->timing;
    try {
      AdvectLaunch L{};
      L.f = &field->view;
      L.seeds = seeds;
      L.rem_in = remaining;
      L.uniform_rem = uniform_remaining;
      L.pos_out = positions_out;
      L.rem_out = remaining_out;
      L.status_out = status_out;
      L.n = count;
      L.dt = step_size;
      L.t0 = start_time;
      L.integrator = integrator;
      L.persistent = true;
      const Tuning& T = tuning_for(ctx, launch_key(field->view, integrator, kSeedArray, false));
      OrderScratch probe;
      L.order_scratch = &probe;   // asks whether this launch orders its queue at all
      const bool ordered = ordered_launch(ctx, T, L);
      L.order_scratch = nullptr;
      if (ordered) {
        uint32_t shift, bits, keys, tiles;
        order_shape(field->view, (uint32_t)T.order_brick, shift, bits, keys, tiles);
        p->key = DeviceBuffer(ctx, count * sizeof(uint32_t));
        p->order = DeviceBuffer(ctx, count * sizeof(uint32_t));
        p->count = DeviceBuffer(ctx, ((uint64_t)keys + tiles) * sizeof(uint32_t));
        p->scratch = OrderScratch{p->key.as<uint32_t>(), p->order.as<uint32_t>(),
                                  p->count.as<uint32_t>(), count, (uint64_t)keys + tiles};
        L.order_scratch = &p->scratch;
      }
      ctx->sync();   // nothing of the caller's in flight on the stream while capturing
      ctx->timing = false;   // no event records inside the graph
      DPAH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      try {
        launch_advect(ctx, L);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      DPAH_HIP(hipStreamEndCapture(ctx->stream, &p->graph));
      ctx->timing = timing;
      DPAH_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    } catch (...) {
      ctx->timing = timing;
      destroy_pass(p);
      throw;
    }
    *out = p;
  });
}

DPAH_EXPORT int dpah_pass_launch(dpah_pass* p) {
  return guarded([&] {
    DPAH_REQUIRE(p, "null argument");
    dpah_context* ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ctx->timed("advect_pass_graph", [&] { DPAH_HIP(hipGraphLaunch(p->exec, ctx->stream)); });
    p->launches += 1;
  });
}

DPAH_EXPORT int dpah_pass_total_steps(dpah_pass* p, uint64_t* steps) {
  return guarded([&] {
    DPAH_REQUIRE(p && steps, "null argument");
    dpah_context* ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    unsigned long long s = 0;
    DPAH_HIP(hipMemcpyAsync(&s, ctx->d_counters + 1, sizeof(s), hipMemcpyDeviceToHost,
                            ctx->stream));
    ctx->sync();
    *steps = s;
  });
}

DPAH_EXPORT int dpah_pass_destroy(dpah_pass* p) {
  return guarded([&] {
    if (!p) return;
    dpah_context* ctx = p->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    destroy_pass(p);
  });
}
