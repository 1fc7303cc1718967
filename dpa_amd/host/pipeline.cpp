// pipeline.cpp -- one rank of dpa's pipeline (source/pipeline.cpp:22-176) on
// one GPU: configuration -> HDF5 field -> seeds -> advection (all rounds) ->
// curves (indices and colours from the GPU) -> HDF5 + XDMF -> FTLE -> CSV.
#include <chrono>
#include <iostream>
#include <limits>
#include <random>

#include "dpa_host.hpp"

namespace dpa_hip {

namespace {
double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

int pipeline::run_config(const arguments& a, int device_ordinal) {
  session_recorder recorder;
  std::cout << "Started pipeline on one GPU\n";

  std::cout << "data_loading\n";
  regular_grid_loader loader(a.input_dataset_filepath, a.input_dataset_name,
                             a.input_dataset_spacing_name);
  regular_vector_field_3d field = loader.load_vector_field();
  device dev(device_ordinal);
  resident_field resident(dev, field);

  std::cout << "seed_generation\n";
  // pipeline.cpp:59-60: offset = spacing * 0, size = spacing * block size
  vector3 offset{}, size{};
  for (int i = 0; i < 3; ++i) {
    offset[i] = field.spacing[i] * 0.0f;
    size[i] = field.spacing[i] * static_cast<float>(field.shape[i]);
  }
  std::vector<particle_3d> particles;
  if (a.seed_generation_stride) {
    vector3 stride;
    for (int i = 0; i < 3; ++i) stride[i] = field.spacing[i] * (*a.seed_generation_stride)[i];
    particles = uniform_seed_generator::generate(offset, size, stride, a.seed_generation_iterations,
                                                 0, a.seed_generation_boundaries);
  } else if (a.seed_generation_count) {
    particles = uniform_seed_generator::generate_random(
        offset, size, *a.seed_generation_count, a.seed_generation_iterations, 0,
        a.seed_generation_boundaries, a.seed_generation_random_seed);
  } else if (a.seed_generation_range) {
    std::mt19937_64 rng(a.seed_generation_random_seed);
    std::uniform_int_distribution<std::size_t> d((*a.seed_generation_range)[0],
                                          (*a.seed_generation_range)[1]);
    particles = uniform_seed_generator::generate_random(
        offset, size, d(rng), a.seed_generation_iterations, 0, a.seed_generation_boundaries,
        a.seed_generation_random_seed);
  }

  std::cout << "particle_advection\n";
  // pipeline.cpp:155: 64-bit indices iff particles_per_round * iterations > 2^32 - 1
  // (size_t product, wrapping like dpa's)
  const bool use_64 = static_cast<std::size_t>(a.particle_advector_particles_per_round *
                                        a.seed_generation_iterations) >
                      std::numeric_limits<std::uint32_t>::max();
  particle_advector advector(dev, a.particle_advector_particles_per_round,
                             a.particle_advector_load_balancer, a.particle_advector_integrator,
                             a.particle_advector_step_size, a.particle_advector_gather_particles,
                             a.particle_advector_record, use_64);
  // total_time covers the round loop only (pipeline.cpp:98-145); the records
  // of each round follow dpa's names and order
  std::vector<particle_advector::round_record> rounds;
  const auto t_loop = std::chrono::steady_clock::now();
  particle_advector::output output = advector.advect(resident, particles, &rounds);
  const double loop_ms = ms_since(t_loop);
  for (std::size_t r = 0; r < rounds.size(); ++r) {
    const std::string k = "round." + std::to_string(r) + ".";
    recorder.set(k + "load", rounds[r].load);
    recorder.set(k + "load_balancing_time", rounds[r].load_balancing_ms);
    recorder.set(k + "advection_time", rounds[r].advection_ms);
    recorder.set(k + "communication_time", rounds[r].communication_ms);
    recorder.set(k + "time", rounds[r].time_ms);
  }
  recorder.set("total_time", loop_ms);

  std::cout << "index_generation\n";
  if (a.particle_advector_record) index_generator::generate(output.integral_curves, use_64);
  std::cout << "color_generation\n";
  if (a.particle_advector_record)
    color_generator::generate_from_angular_velocities(output.integral_curves);
  std::cout << "save_integral_curves\n";
  if (a.particle_advector_record)
    integral_curve_saver(a.output_dataset_filepath).save(output.integral_curves);

  std::cout << "estimate_ftle\n";
  if (a.estimate_ftle) {
    if (!a.seed_generation_stride)
      throw error(DPAH_ERR_INVALID, "estimate_ftle needs seed_generation_stride");
    regular_scalar_field_3d ftle = ftle_estimator::estimate(
        dev, field, a.seed_generation_iterations, *a.seed_generation_stride,
        a.particle_advector_step_size, output.inactive_particles);
    std::cout << "save_ftle_field\n";
    regular_grid_saver(a.output_dataset_filepath).save(ftle);
  }
  recorder.to_csv(a.output_dataset_filepath + ".benchmark.csv");
  return 0;
}

int pipeline::run(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: " << argv[0] << " <config.json>\n";
    return 2;
  }
  return run_config(argument_parser::parse(argv[1]));
}

}  // namespace dpa_hip
