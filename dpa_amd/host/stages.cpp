// stages.cpp -- the stage classes of dpa_host.hpp over the C ABI.
#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <random>

#include "dpa_host.hpp"

namespace dpa_hip {

void check(int code) {
  if (code != DPAH_OK) throw error(code, dpah_last_error());
}

device::device(int ordinal) { check(dpah_create(ordinal, nullptr, &ctx_)); }
device::~device() { dpah_destroy(ctx_); }

resident_field::resident_field(device& dev, const regular_vector_field_3d& f) : host_(f) {
  const uint64_t shape[3] = {f.shape[0], f.shape[1], f.shape[2]};
  check(dpah_field_create(dev.get(), f.data.data(), shape, f.offset.data(), f.spacing.data(),
                          &field_));
}
resident_field::~resident_field() { dpah_field_release(field_); }

// particle_advector.cpp:29-36: dpa maps the integrator name to an odeint
// stepper and silently falls back to Euler for unknown names.  The GPU path
// implements runge_kutta_4 (dpa's one-sample form) and classic four-sample
// RK4; other names are rejected instead of running a different scheme.
static int integrator_code(const std::string& name) {
  if (name == "runge_kutta_4") return DPAH_INTEGRATOR_DPA_RK4;
  if (name == "runge_kutta_4_classic") return DPAH_INTEGRATOR_RK4;
  throw error(DPAH_ERR_UNSUPPORTED, "integrator '" + name +
                                        "' is not available on the GPU path "
                                        "(runge_kutta_4, runge_kutta_4_classic)");
}

particle_advector::particle_advector(device& dev, std::size_t particles_per_round,
                                     const std::string& /*load_balancer*/,
                                     const std::string& integrator, scalar step_size,
                                     bool /*gather_particles*/, bool record,
                                     bool use_64_bit_indices)
    : dev_(dev),
      particles_per_round_(particles_per_round),
      integrator_(integrator_code(integrator)),
      step_size_(step_size),
      record_(record),
      use_64_bit_(use_64_bit_indices) {}

particle_advector::output particle_advector::advect(const resident_field& field,
                                                    std::vector<particle_3d>& particles,
                                                    std::vector<round_record>* rounds) {
  using clock = std::chrono::steady_clock;
  auto ms = [](clock::time_point a, clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  const std::size_t n = particles.size();
  output out;
  out.inactive_particles.resize(n);
  std::size_t active = n;
  do {   // pipeline.cpp:100-146: at least one round, the tail of the particles first
    round_record rec;
    const auto t0 = clock::now();
    rec.load = static_cast<double>(active);   // no load balancer: every active particle
    const auto t1 = clock::now();
    // compute_round_state, allocate / advect / prune (particle_advector.cpp:261-408):
    // one call over the round's slice, which is one round of the library's loop
    const std::size_t count = std::min(particles_per_round_, active);
    const std::size_t first = active - count;
    std::vector<float> seeds(3 * count);
    std::vector<uint64_t> remaining(count);
    for (std::size_t i = 0; i < count; ++i) {
      std::memcpy(&seeds[3 * i], particles[first + i].position.data(), 3 * sizeof(float));
      remaining[i] = particles[first + i].remaining_iterations;
    }
    dpah_advect_params p{};
    p.particles_per_round = count ? count : 1;
    p.step_size = step_size_;
    p.integrator = integrator_;
    p.record = record_ ? 1 : 0;
    p.use_64_bit_indices = use_64_bit_ ? 1 : 0;
    dpah_result* r = nullptr;
    check(dpah_advect(dev_.get(), field.get(), seeds.data(), remaining.data(), 0, count, &p, &r));
    try {
      if (record_) {
        uint64_t V = 0, I = 0;
        int ib = 4;
        check(dpah_result_round_sizes(r, 0, &V, &I, &ib));
        integral_curve c;
        c.vertices.resize(V);
        std::vector<scalar> colors(V);
        if (ib == 8) {
          std::vector<uint64_t> idx(I);
          check(dpah_result_copy_round(r, 0, c.vertices.data()->data(), colors.data(), idx.data()));
          c.indices = std::move(idx);
        } else {
          std::vector<uint32_t> idx(I);
          check(dpah_result_copy_round(r, 0, c.vertices.data()->data(), colors.data(), idx.data()));
          c.indices = std::move(idx);
        }
        c.colors = std::move(colors);
        out.integral_curves.push_back(std::move(c));
      }
      std::vector<float> pos(3 * count);
      std::vector<uint64_t> rem(count);
      std::vector<uint8_t> status(count);
      check(dpah_result_final_state(r, pos.data(), rem.data(), status.data()));
      for (std::size_t i = 0; i < count; ++i) {
        particle_3d q = particles[first + i];
        std::memcpy(q.position.data(), &pos[3 * i], 3 * sizeof(float));
        q.remaining_iterations = rem[i];
        out.inactive_particles[first + i] = q;
      }
    } catch (...) {
      dpah_result_free(r);
      throw;
    }
    dpah_result_free(r);
    const auto t2 = clock::now();
    active = first;   // completion check (particle_advector.cpp:57-64): one rank
    const auto t3 = clock::now();
    rec.load_balancing_ms = ms(t0, t1);
    rec.advection_ms = ms(t1, t2);
    rec.communication_ms = ms(t2, t3);
    rec.time_ms = ms(t0, t3);
    if (rounds) rounds->push_back(rec);
  } while (active > 0);
  particles.clear();   // advect consumes the active particles (particle_advector.cpp:398)
  return out;
}

void index_generator::generate(integral_curves& curves, bool use_64_bit) {
  for (auto& c : curves) {
    if (c.vertices.empty()) continue;
    if (use_64_bit && std::holds_alternative<std::vector<uint32_t>>(c.indices)) {
      const auto& src = std::get<std::vector<uint32_t>>(c.indices);
      std::vector<uint64_t> dst(src.begin(), src.end());
      c.indices = std::move(dst);
    } else if (!use_64_bit && std::holds_alternative<std::vector<uint64_t>>(c.indices)) {
      const auto& src = std::get<std::vector<uint64_t>>(c.indices);
      if (c.vertices.size() > 0xFFFFFFFFull)
        throw error(DPAH_ERR_INVALID, "32-bit indices cannot address this round");
      std::vector<uint32_t> dst(src.begin(), src.end());
      c.indices = std::move(dst);
    }
  }
}

void color_generator::generate_from_angular_velocities(integral_curves& curves) {
  for (auto& c : curves)
    if (!std::holds_alternative<std::vector<scalar>>(c.colors) ||
        std::get<std::vector<scalar>>(c.colors).size() != c.vertices.size())
      throw error(DPAH_ERR_INVALID, "curves without GPU colours");
}

regular_scalar_field_3d ftle_estimator::estimate(device& dev, const regular_vector_field_3d& f,
                                                 std::size_t iterations, const vector3& stride,
                                                 scalar step_size,
                                                 const std::vector<particle_3d>& particles) {
  const std::size_t n = particles.size();
  std::vector<float> orig(3 * n), fin(3 * n);
  for (std::size_t i = 0; i < n; ++i) {
    std::memcpy(&orig[3 * i], particles[i].original_position.data(), 12);
    std::memcpy(&fin[3 * i], particles[i].position.data(), 12);
  }
  dpah_context* ctx = dev.get();
  void *d_o = nullptr, *d_f = nullptr, *d_out = nullptr;
  const uint64_t shape[3] = {f.shape[0], f.shape[1], f.shape[2]};
  uint64_t m[3] = {0, 0, 0};
  check(dpah_ftle_estimate_device(ctx, shape, f.offset.data(), f.spacing.data(), stride.data(),
                                  iterations, step_size, nullptr, nullptr, n, nullptr, m));
  regular_scalar_field_3d out;
  out.shape = {m[0], m[1], m[2]};
  out.offset = f.offset;
  out.size = f.size;
  for (int i = 0; i < 3; ++i) out.spacing[i] = f.spacing[i] * stride[i];
  out.data.resize(m[0] * m[1] * m[2]);
  try {
    check(dpah_device_alloc(ctx, std::max<std::size_t>(orig.size(), 1) * 4, &d_o));
    check(dpah_device_alloc(ctx, std::max<std::size_t>(fin.size(), 1) * 4, &d_f));
    check(dpah_device_alloc(ctx, std::max<std::size_t>(out.data.size(), 1) * 4, &d_out));
    check(dpah_copy_to_device(ctx, d_o, orig.data(), orig.size() * 4));
    check(dpah_copy_to_device(ctx, d_f, fin.data(), fin.size() * 4));
    check(dpah_ftle_estimate_device(ctx, shape, f.offset.data(), f.spacing.data(), stride.data(),
                                    iterations, step_size, static_cast<float*>(d_o),
                                    static_cast<float*>(d_f), n, static_cast<float*>(d_out), m));
    check(dpah_copy_to_host(ctx, out.data.data(), d_out, out.data.size() * 4));
  } catch (...) {
    dpah_device_free(ctx, d_o);
    dpah_device_free(ctx, d_f);
    dpah_device_free(ctx, d_out);
    throw;
  }
  dpah_device_free(ctx, d_o);
  dpah_device_free(ctx, d_f);
  dpah_device_free(ctx, d_out);
  return out;
}

// uniform_seed_generator.cpp:18-28: restrict to the box when it intersects.
static bool clip(vector3& offset, vector3& size, const std::optional<aabb3>& box) {
  if (!box) return true;
  vector3 lo, hi;
  for (int i = 0; i < 3; ++i) {
    lo[i] = std::max(box->minimum[i], offset[i]);
    const float end = offset[i] + size[i];
    hi[i] = std::min(box->maximum[i], end);
    if (hi[i] < lo[i]) return false;   // Eigen AlignedBox::isEmpty
  }
  for (int i = 0; i < 3; ++i) {
    offset[i] = lo[i];
    size[i] = hi[i] - lo[i];
  }
  return true;
}

std::vector<particle_3d> uniform_seed_generator::generate(vector3 offset, vector3 size,
                                                          vector3 stride, std::size_t iterations,
                                                          integer process_index,
                                                          std::optional<aabb3> aabb) {
  if (!clip(offset, size, aabb)) return {};
  std::array<std::size_t, 3> per{};
  for (int i = 0; i < 3; ++i) {
    const float q = size[i] / stride[i];
    per[i] = q > 0.0f ? static_cast<std::size_t>(q) : 0;   // (size / stride).cast<size_t>()
  }
  const std::size_t count = per[0] * per[1] * per[2];
  std::vector<particle_3d> out(count);
  for (std::size_t index = 0; index < count; ++index) {
    std::size_t rest = index, sub[3];
    for (int i = 2; i >= 0; --i) {                      // indexing.hpp:10-20, C order
      sub[i] = rest % per[i];
      rest /= per[i];
    }
    particle_3d& p = out[index];
    for (int i = 0; i < 3; ++i) {
      const float m = stride[i] * static_cast<float>(sub[i]);
      p.position[i] = offset[i] + m;
    }
    p.remaining_iterations = iterations;
    p.original_rank = process_index;
    p.original_position = p.position;
  }
  return out;
}

std::vector<particle_3d> uniform_seed_generator::generate_random(
    vector3 offset, vector3 size, std::size_t count, std::size_t iterations,
    integer process_index, std::optional<aabb3> aabb, std::uint64_t random_seed) {
  if (!clip(offset, size, aabb)) return {};
  std::mt19937_64 rng(random_seed);
  std::vector<particle_3d> out(count);
  for (auto& p : out) {
    for (int i = 0; i < 3; ++i) {
      std::uniform_real_distribution<float> d(offset[i], offset[i] + size[i]);
      p.position[i] = d(rng);
    }
    p.remaining_iterations = iterations;
    p.original_rank = process_index;
    p.original_position = p.position;
  }
  return out;
}

}  // namespace dpa_hip
