// dpa_host.hpp -- host C++ above the C ABI, mirroring the stage classes of one
// dpa rank (names, arguments and meaning of include/dpa/stages/*.hpp and
// include/dpa/types/*.hpp) so that dpa's single-rank pipeline runs unchanged
// on one MI355X.  Every computation of the advection path happens behind
// include/dpa_hip.h; this layer moves data and files.
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "../../include/dpa_hip.h"

namespace dpa_hip {

using scalar = float;
using integer = std::int32_t;
using vector3 = std::array<scalar, 3>;

// basic_types.hpp:48 (Eigen::AlignedBox<float,3>)
struct aabb3 {
  vector3 minimum{}, maximum{};
};

// relative_direction.hpp:6-15; one GPU holds the whole block, so every
// particle is `center`.
enum relative_direction { center = 0 };

// particle.hpp:65-72 with DPA_FTLE_SUPPORT
struct particle_3d {
  vector3 position{};
  std::size_t remaining_iterations = 0;
  relative_direction direction = center;
  integer original_rank = 0;
  vector3 original_position{};
};

// integral_curves.hpp:11-18
struct integral_curve {
  std::vector<vector3> vertices;
  std::variant<std::vector<scalar>, std::vector<std::array<std::uint8_t, 3>>> colors;
  std::variant<std::vector<std::uint32_t>, std::vector<std::uint64_t>> indices;
};
using integral_curves = std::vector<integral_curve>;

// regular_grid<vector3,3> / regular_grid<float,3> (regular_grid.hpp:148-151):
// C-order data, offset, size and spacing in physical units.
struct regular_vector_field_3d {
  std::vector<float> data;            // [n0][n1][n2][3]
  std::array<std::size_t, 3> shape{};
  vector3 offset{}, size{}, spacing{};
};
struct regular_scalar_field_3d {
  std::vector<float> data;            // [n0][n1][n2]
  std::array<std::size_t, 3> shape{};
  vector3 offset{}, size{}, spacing{};
};

// arguments.hpp:11-29
struct arguments {
  std::string input_dataset_filepath;
  std::string input_dataset_name;
  std::string input_dataset_spacing_name;
  std::optional<vector3> seed_generation_stride;
  std::optional<std::size_t> seed_generation_count;
  std::optional<std::array<std::size_t, 2>> seed_generation_range;
  std::size_t seed_generation_iterations = 0;
  std::optional<aabb3> seed_generation_boundaries;
  std::size_t particle_advector_particles_per_round = 0;
  std::string particle_advector_load_balancer;
  std::string particle_advector_integrator;
  scalar particle_advector_step_size = 0;
  bool particle_advector_gather_particles = false;
  bool particle_advector_record = false;
  bool estimate_ftle = false;
  std::string output_dataset_filepath;
  std::uint64_t seed_generation_random_seed = 0;   // extension: random seeding is reproducible
};

struct error : std::runtime_error {
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void check(int code);   // throws dpa_hip::error with dpah_last_error() unless DPAH_OK

// The GPU (a dpah_context).
class device {
 public:
  explicit device(int ordinal = 0);
  ~device();
  device(const device&) = delete;
  device& operator=(const device&) = delete;
  dpah_context* get() const { return ctx_; }

 private:
  dpah_context* ctx_ = nullptr;
};

// A field resident in HBM (what regular_grid_loader's result becomes on the GPU).
class resident_field {
 public:
  resident_field(device& dev, const regular_vector_field_3d& field);
  ~resident_field();
  resident_field(const resident_field&) = delete;
  resident_field& operator=(const resident_field&) = delete;
  dpah_field* get() const { return field_; }
  const regular_vector_field_3d& host() const { return host_; }

 private:
  dpah_field* field_ = nullptr;
  const regular_vector_field_3d& host_;
};

// particle_advector.hpp:22-160 for one rank.  The load balancer and the
// neighbour exchange have nothing to do on one GPU; gather_particles is a
// no-op for the same reason.
class particle_advector {
 public:
  struct output {
    std::vector<particle_3d> inactive_particles;
    dpa_hip::integral_curves integral_curves;
  };
  // One round of dpa's loop (pipeline.cpp:100-146): the active particles at
  // its start ("load") and the wall time of its stages in milliseconds.
  struct round_record {
    double load = 0;
    double load_balancing_ms = 0, advection_ms = 0, communication_ms = 0, time_ms = 0;
  };

  particle_advector(device& dev, std::size_t particles_per_round, const std::string& load_balancer,
                    const std::string& integrator, scalar step_size, bool gather_particles,
                    bool record, bool use_64_bit_indices = false);

  // Advects and consumes `particles` (particle_advector.cpp:39-55, all rounds),
  // with indices and colours already generated.  The rounds run as dpa's loop
  // does, the tail of the active particles first, one library call each; a
  // record per round goes to `rounds` when it is given.  inactive_particles
  // keeps the seed order.
  output advect(const resident_field& field, std::vector<particle_3d>& particles,
                std::vector<round_record>* rounds = nullptr);

 private:
  device& dev_;
  std::size_t particles_per_round_;
  int integrator_;
  scalar step_size_;
  bool record_;
  bool use_64_bit_;
};

// index_generator.hpp:8-12.  The GPU writes the indices with the curves; this
// only converts them when the requested width differs.
struct index_generator {
  static void generate(integral_curves& curves, bool use_64_bit = false);
};

// color_generator.hpp:10-17.  Angular-velocity colours come with the curves.
struct color_generator {
  static void generate_from_angular_velocities(integral_curves& curves);
};

// ftle_estimator.hpp:11-20
struct ftle_estimator {
  static regular_scalar_field_3d estimate(device& dev, const regular_vector_field_3d& field,
                                          std::size_t iterations, const vector3& stride,
                                          scalar step_size,
                                          const std::vector<particle_3d>& particles);
};

// uniform_seed_generator.hpp:12-22
struct uniform_seed_generator {
  static std::vector<particle_3d> generate(vector3 offset, vector3 size, vector3 stride,
                                           std::size_t iterations, integer process_index,
                                           std::optional<aabb3> aabb);
  // dpa draws from std::random_device per thread, so its random seeds cannot be
  // reproduced; this one uses mt19937_64 with an explicit seed.
  static std::vector<particle_3d> generate_random(vector3 offset, vector3 size,
                                                  std::size_t count, std::size_t iterations,
                                                  integer process_index,
                                                  std::optional<aabb3> aabb,
                                                  std::uint64_t random_seed);
};

// argument_parser.hpp
struct argument_parser {
  static arguments parse(const std::string& filepath);
  static arguments parse_text(const std::string& text);
};

// regular_grid_loader.hpp:16-36 (serial HDF5; one rank reads the whole grid)
struct regular_grid_loader {
  regular_grid_loader(std::string filepath, std::string dataset, std::string spacing);
  std::array<std::size_t, 3> load_dimensions() const;
  regular_vector_field_3d load_vector_field() const;
  std::string filepath, dataset, spacing;
};

// integral_curve_saver.hpp / regular_grid_saver.hpp (rank 0)
struct integral_curve_saver {
  explicit integral_curve_saver(const std::string& filepath, int rank = 0);
  void save(const integral_curves& curves) const;
  std::string filepath;   // <stem>.rank_<r>.h5
};
struct regular_grid_saver {
  explicit regular_grid_saver(const std::string& filepath, int rank = 0);
  void save(const regular_scalar_field_3d& field) const;
  std::string filepath;   // <stem>.rank_<r>_grid.h5
};

// HDF5 helpers shared by the CLI (field files in dpa's layout).
void write_vector_field_h5(const std::string& path, const std::string& dataset,
                           const std::string& spacing_attribute,
                           const regular_vector_field_3d& field);
bool hdf5_available();

// benchmark.hpp session CSV ("rank,name,iteration 0,")
struct session_recorder {
  void set(const std::string& name, double value);
  void to_csv(const std::string& filepath, int rank = 0) const;
  std::vector<std::pair<std::string, double>> records;
};

// pipeline.hpp:8-12 / pipeline.cpp:22-176 for one rank on one GPU
struct pipeline {
  static int run(int argc, char** argv);
  static int run_config(const arguments& a, int device_ordinal = 0);
};

std::string to_json(const arguments& a);

}  // namespace dpa_hip
