// ftle.hip -- FTLE entry points: dpa's lattice estimator and the
// auxiliary-neighbour form (ftle_kernels.hpp).
#include <algorithm>
#include <memory>
#include <cmath>
#include <vector>

#include "ftle_kernels.hpp"
#include "runtime.hpp"

using namespace dpah;

namespace {

// |step_size * iterations| in float, as ftle_estimator.cpp:57 computes it.
float ftle_time(float dt, uint64_t iterations) {
  const float T = dt * (float)iterations;
  return std::fabs(T);
}

unsigned grid_for(uint64_t n, unsigned block = 256) {
  return (unsigned)((n + block - 1) / block);
}

// Strided flow-map shape of ftle_estimator.cpp:16-21.
void strided_shape(const uint64_t n[3], const float stride[3], uint64_t m[3]) {
  for (int i = 0; i < 3; ++i) {
    const float q = (float)n[i] / stride[i];
    m[i] = q > 0.0f ? (uint64_t)q : 0;
  }
}

// ftle_estimator::estimate on device data.  `original` null means the
// particles are the lattice `lat` (their original positions are recomputed).
void estimate(dpah_context* ctx, const uint64_t field_n[3], const float offset[3],
              const float spacing[3], const float stride[3], uint64_t iterations, float dt,
              const float* original, const SeedLattice* lat, const float* finals, uint64_t count,
              float* out, uint64_t m[3], uint64_t* misses_out) {
  strided_shape(field_n, stride, m);
  const uint64_t cells = m[0] * m[1] * m[2];
  if (!out || cells == 0) return;
  DeviceBuffer owner(ctx, cells * sizeof(unsigned long long));
  DeviceBuffer flow(ctx, cells * 3 * sizeof(float));
  FtleScatterArgs a{};
  a.original = original;
  if (lat) a.lat = *lat;
  a.finals = finals;
  a.count = count;
  a.o0 = offset[0]; a.o1 = offset[1]; a.o2 = offset[2];
  a.s0 = spacing[0] * stride[0]; a.s1 = spacing[1] * stride[1]; a.s2 = spacing[2] * stride[2];
  a.m0 = m[0]; a.m1 = m[1]; a.m2 = m[2];
  a.owner = owner.as<unsigned long long>();
  a.flow = flow.as<float>();
  a.misses = ctx->d_counters + 2;
  a.T_abs = ftle_time(dt, iterations);
  a.out = out;
  hipLaunchKernelGGL(k_ftle_clear, dim3(grid_for(std::max<uint64_t>(cells, 1))), dim3(256), 0,
                     ctx->stream, a);
  if (count) {
    hipLaunchKernelGGL(k_ftle_claim, dim3(grid_for(count)), dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_ftle_scatter, dim3(grid_for(count)), dim3(256), 0, ctx->stream, a);
  }
  ctx->timed("ftle_gradient", [&] {
    hipLaunchKernelGGL(k_ftle_gradient, dim3(grid_for(cells)), dim3(256), 0, ctx->stream, a);
  });
  DPAH_HIP(hipGetLastError());
  unsigned long long misses = 0;
  DPAH_HIP(hipMemcpyAsync(&misses, ctx->d_counters + 2, sizeof(misses), hipMemcpyDeviceToHost,
                          ctx->stream));
  ctx->sync();
  if (misses_out) *misses_out = misses;
}

// uniform_seed_generator::generate's lattice (uniform_seed_generator.cpp:16-45)
// for offset/size/stride in physical units.
SeedLattice seed_lattice(const float offset[3], const float size[3], const float stride[3],
                         uint64_t per_dim[3]) {
  SeedLattice L{};
  for (int i = 0; i < 3; ++i) {
    const float q = size[i] / stride[i];
    per_dim[i] = q > 0.0f ? (uint64_t)q : 0;
  }
  L.lo0 = offset[0]; L.lo1 = offset[1]; L.lo2 = offset[2];
  L.st0 = stride[0]; L.st1 = stride[1]; L.st2 = stride[2];
  L.d0 = per_dim[0]; L.d1 = per_dim[1]; L.d2 = per_dim[2];
  return L;
}

}  // namespace

DPAH_EXPORT int dpah_ftle_estimate_device(dpah_context* ctx, const uint64_t field_shape[3],
                                          const float offset[3], const float spacing[3],
                                          const float stride[3], uint64_t iterations,
                                          float step_size, const float* original_positions,
                                          const float* final_positions, uint64_t count,
                                          float* ftle_out, uint64_t out_shape[3]) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field_shape && offset && spacing && stride && out_shape, "null argument");
    DPAH_REQUIRE(count == 0 || !ftle_out || (original_positions && final_positions),
                 "positions are null");
    for (int i = 0; i < 3; ++i) DPAH_REQUIRE(stride[i] > 0.0f, "stride must be positive");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    estimate(ctx, field_shape, offset, spacing, stride, iterations, step_size, original_positions,
             nullptr, final_positions, count, ftle_out, out_shape, nullptr);
  });
}

DPAH_EXPORT int dpah_ftle_lattice(dpah_context* ctx, const dpah_field* field, const float stride[3],
                                  uint64_t iterations, float step_size, int integrator,
                                  float* ftle_out, uint64_t out_shape[3], uint64_t* total_steps) {
  return guarded([&] {
    DPAH_REQUIRE(ctx && field && stride && out_shape, "null argument");
    DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
    DPAH_REQUIRE(field->time_steps == 1, "dpa's FTLE stage takes a steady field");
    DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
                 "unknown integrator");
    DPAH_REQUIRE(iterations < (1ull << 32) - 2, "iterations must be below 2^32 - 2");
    for (int i = 0; i < 3; ++i) DPAH_REQUIRE(stride[i] > 0.0f, "stride must be positive");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    const FieldView& v = field->view;
    const float offset[3] = {v.o0, v.o1, v.o2};
    const float spacing[3] = {v.h0, v.h1, v.h2};
    // pipeline.cpp:59-72 on one rank: offset = spacing * 0, size = spacing * n,
    // stride in cells times the spacing
    float size[3], st[3];
    for (int i = 0; i < 3; ++i) {
      size[i] = spacing[i] * (float)field->shape[i];
      st[i] = spacing[i] * stride[i];
    }
    uint64_t per[3];
    const SeedLattice lat = seed_lattice(offset, size, st, per);
    const uint64_t count = per[0] * per[1] * per[2];
    strided_shape(field->shape, stride, out_shape);
    if (!ftle_out) return;
    DeviceBuffer finals(ctx, std::max<uint64_t>(count, 1) * 3 * sizeof(float));
    AdvectLaunch L{};
    L.f = &v;
    L.seed_kind = kSeedLattice;
    L.lat = lat;
    L.uniform_rem = iterations;
    L.pos_out = finals.as<float>();
    L.n = count;
    L.dt = step_size;
    L.integrator = integrator;
    L.persistent = true;
    launch_advect(ctx, L);
    unsigned long long s = 0;
    DPAH_HIP(hipMemcpyAsync(&s, ctx->d_counters + 1, sizeof(s), hipMemcpyDeviceToHost,
                            ctx->stream));
    ctx->sync();
    if (total_steps) *total_steps = s;
    const uint64_t cells = out_shape[0] * out_shape[1] * out_shape[2];
    DeviceBuffer out(ctx, std::max<uint64_t>(cells, 1) * sizeof(float));
    uint64_t m[3];
    estimate(ctx, field->shape, offset, spacing, stride, iterations, step_size, nullptr, &lat,
             finals.as<float>(), count, out.as<float>(), m, nullptr);
    ctx->download(ftle_out, out.ptr, cells * sizeof(float), ctx->stream);
    ctx->sync();
  });
}

int dpah::aux_lattice(SeedLattice& lat, const float origin[3], const float lattice_spacing[3],
                      const uint64_t shape[3], const float delta[3], uint64_t* particles) {
  lat = SeedLattice{};
  lat.lo0 = origin[0]; lat.lo1 = origin[1]; lat.lo2 = origin[2];
  lat.st0 = lattice_spacing[0]; lat.st1 = lattice_spacing[1]; lat.st2 = lattice_spacing[2];
  lat.d0 = shape[0]; lat.d1 = shape[1]; lat.d2 = shape[2];
  lat.dl0 = delta[0]; lat.dl1 = delta[1]; lat.dl2 = delta[2];
  bool shared = true;
  for (int i = 0; i < 3; ++i)   // st = 2 delta exactly, so st - delta = delta
    shared = shared && delta[i] > 0.0f && 2.0f * delta[i] == lattice_spacing[i] &&
             shape[i] < (1ull << 24);
  const uint64_t points = shape[0] * shape[1] * shape[2];
  if (particles)   // (an empty lattice has no face particles either)
    *particles = points == 0 ? 0 : shared ? aux_shared_count(shape[0], shape[1], shape[2])
                                          : 6 * points;
  return shared ? kSeedAuxShared : kSeedAux;
}

void dpah::aux_reduce(const float* finals, int kind, const SeedLattice& lat, const float delta[3],
                      float step_size, uint64_t iterations, float* out, hipStream_t s) {
  FtleAuxArgs a{};
  a.finals = finals;
  a.points = lat.d0 * lat.d1 * lat.d2;
  a.twod0 = 2.0f * delta[0]; a.twod1 = 2.0f * delta[1]; a.twod2 = 2.0f * delta[2];
  a.T_abs = ftle_time(step_size, iterations);
  a.out = out;
  a.shared = kind == kSeedAuxShared ? 1u : 0u;
  a.d0 = lat.d0; a.d1 = lat.d1; a.d2 = lat.d2;
  if (a.points)
    hipLaunchKernelGGL(k_ftle_aux_reduce, dim3(grid_for(a.points)), dim3(256), 0, s, a);
  DPAH_HIP(hipGetLastError());
}

void dpah::ftle_aux_into(dpah_context* ctx, const dpah_field* field, const float origin[3],
                         const float lattice_spacing[3], const uint64_t lattice_shape[3],
                         const float delta[3], uint64_t iterations, float step_size,
                         float start_time, int integrator, float* ftle_out, bool persistent,
                         uint64_t* total_steps) {
  DPAH_REQUIRE(ctx && field && origin && lattice_spacing && lattice_shape && delta,
               "null argument");
  DPAH_REQUIRE(field->ctx == ctx, "field belongs to another context");
  DPAH_REQUIRE(integrator == DPAH_INTEGRATOR_DPA_RK4 || integrator == DPAH_INTEGRATOR_RK4,
               "unknown integrator");
  DPAH_REQUIRE(iterations < (1ull << 32) - 2, "iterations must be below 2^32 - 2");
  const uint64_t points = lattice_shape[0] * lattice_shape[1] * lattice_shape[2];
  DPAH_REQUIRE(points == 0 || ftle_out, "ftle_out is null");
  if (points == 0) {   // an empty lattice (some axis of no points): nothing to advect
    if (total_steps) *total_steps = 0;
    return;
  }
  SeedLattice lat;
  const int kind = aux_lattice(lat, origin, lattice_spacing, lattice_shape, delta, nullptr);
  // points are processed in chunks of at most 2^25 whose final positions take
  // at most a quarter of the memory budget; with shared particles in whole
  // slabs along axis 0 (each chunk advects its own low-face particles, so a
  // chunk after the first repeats one slab of the previous chunk's +delta
  // particles along axis 0, and its particle-steps are counted again)
  const uint64_t plane = lat.d1 * lat.d2;
  const bool shared = kind == kSeedAuxShared && plane <= (1ull << 25);
  const uint64_t per_point = (shared ? 3 : 6) * 3 * sizeof(float);
  uint64_t cap = std::min<uint64_t>(1ull << 25,
                                    std::max<uint64_t>(ctx->memory_budget / 4 / per_point, 1));
  if (shared) cap = std::max<uint64_t>(cap / plane, 1) * plane;
  const uint64_t chunk = std::min<uint64_t>(points, cap);
  const uint64_t per_chunk =
      shared ? aux_shared_count(chunk / plane, lat.d1, lat.d2) : 6 * chunk;
  DeviceBuffer finals(ctx, std::max<uint64_t>(per_chunk, 1) * 3 * sizeof(float));
  uint64_t steps = 0;
  for (uint64_t p0 = 0; p0 < points; p0 += chunk) {
    const uint64_t pn = std::min(chunk, points - p0);
    AdvectLaunch L{};
    L.f = &field->view;
    L.seed_kind = shared ? kSeedAuxShared : kSeedAux;
    L.lat = lat;
    L.lat.first = p0;
    if (shared) L.lat.d0 = pn / plane;   // the chunk's slabs
    L.uniform_rem = iterations;
    L.pos_out = finals.as<float>();
    L.n = shared ? aux_shared_count(L.lat.d0, lat.d1, lat.d2) : 6 * pn;
    L.dt = step_size;
    L.t0 = start_time;
    L.integrator = integrator;
    L.persistent = persistent;
    launch_advect(ctx, L);
    SeedLattice part = L.lat;
    if (!shared) { part.d0 = pn; part.d1 = 1; part.d2 = 1; }   // six per point: points only
    ctx->timed("ftle_aux_reduce", [&] {
      aux_reduce(finals.as<float>(), L.seed_kind, part, delta, step_size, iterations,
                 ftle_out + p0, ctx->stream);
    });
    unsigned long long s = 0;
    DPAH_HIP(hipMemcpyAsync(&s, ctx->d_counters + 1, sizeof(s), hipMemcpyDeviceToHost,
                            ctx->stream));
    ctx->sync();
    steps += s;
  }
  if (total_steps) *total_steps = steps;
}

dpah_result* dpah::run_ftle_aux(dpah_context* ctx, const dpah_field* field, const float origin[3],
                                const float lattice_spacing[3], const uint64_t lattice_shape[3],
                                const float delta[3], uint64_t iterations, float step_size,
                                float start_time, int integrator, bool persistent) {
  auto res = std::make_unique<dpah_result>();
  res->ctx = ctx;
  res->ftle_count = lattice_shape[0] * lattice_shape[1] * lattice_shape[2];
  res->d_ftle = static_cast<float*>(ctx->alloc(std::max<uint64_t>(res->ftle_count, 1) * 4));
  uint64_t steps = 0;
  ftle_aux_into(ctx, field, origin, lattice_spacing, lattice_shape, delta, iterations, step_size,
                start_time, integrator, res->d_ftle, persistent, &steps);
  res->total_steps = steps;
  return res.release();
}

DPAH_EXPORT int dpah_ftle_aux_device(dpah_context* ctx, const dpah_field* field,
                                     const float origin[3], const float lattice_spacing[3],
                                     const uint64_t lattice_shape[3], const float delta[3],
                                     uint64_t iterations, float step_size, float start_time,
                                     int integrator, float* ftle_out, uint64_t* total_steps) {
  return guarded([&] {
    DPAH_REQUIRE(ctx, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DPAH_HIP(hipSetDevice(ctx->device));
    ftle_aux_into(ctx, field, origin, lattice_spacing, lattice_shape, delta, iterations, step_size,
                  start_time, integrator, ftle_out, true, total_steps);
  });
}
