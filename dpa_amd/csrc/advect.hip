// advect.hip -- launchers of the K_advect kernels (advect_kernels.hpp).
#include "advect_kernels.hpp"
#include "runtime.hpp"

namespace dpah {

namespace {

template <bool kRecord, int kIntegrator, bool kUnsteady>
void launch_t(dpah_context* ctx, const AdvectArgs& a, bool persistent) {
  if (a.n == 0) return;
  if (persistent) {
    const int block = ctx->tune.block;
    uint64_t grid = (uint64_t)ctx->cu_count * ctx->tune.blocks_per_cu;
    const uint64_t need = (a.n + block - 1) / block;
    if (grid > need) grid = need;
    ctx->timed("advect_persistent", [&] {
      hipLaunchKernelGGL((k_advect_persistent<kRecord, kIntegrator, kUnsteady>), dim3((unsigned)grid),
                         dim3(block), 0, ctx->stream, a);
    });
  } else {
    const uint64_t grid = (a.n + 255) / 256;
    ctx->timed("advect_simple", [&] {
      hipLaunchKernelGGL((k_advect_simple<kRecord, kIntegrator, kUnsteady>), dim3((unsigned)grid),
                         dim3(256), 0, ctx->stream, a);
    });
  }
  DPAH_HIP(hipGetLastError());
}

template <bool kRecord>
void dispatch_integrator(dpah_context* ctx, const AdvectArgs& a, int integrator, bool unsteady,
                         bool persistent) {
  if (integrator == kRk4) {
    if (unsteady) launch_t<kRecord, kRk4, true>(ctx, a, persistent);
    else          launch_t<kRecord, kRk4, false>(ctx, a, persistent);
  } else {
    if (unsteady) launch_t<kRecord, kDpaRk4, true>(ctx, a, persistent);
    else          launch_t<kRecord, kDpaRk4, false>(ctx, a, persistent);
  }
}

__global__ void k_reset_counters(unsigned long long* c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0ull;
}

}  // namespace

void launch_advect(dpah_context* ctx, const AdvectLaunch& L) {
  AdvectArgs a{};
  a.f = *L.f;
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
  a.queue = ctx->d_counters;
  a.total_steps = ctx->d_counters + 1;
  // queue head and step counter start at zero for every launch (a kernel, not
  // a memset, so the sequence stays valid inside a captured graph)
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, ctx->stream, ctx->d_counters);
  DPAH_HIP(hipGetLastError());
  const bool unsteady = L.f->nt > 1;
  if (L.slots) dispatch_integrator<true>(ctx, a, L.integrator, unsteady, L.persistent);
  else         dispatch_integrator<false>(ctx, a, L.integrator, unsteady, L.persistent);
}

}  // namespace dpah
