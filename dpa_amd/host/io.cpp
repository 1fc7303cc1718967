// io.cpp -- configuration, HDF5 / XDMF files and the timing CSV, in dpa's
// formats (argument_parser.cpp:10-60, regular_grid_loader.cpp:16-88,
// integral_curve_saver.cpp:19-88, regular_grid_saver.cpp:14-46, xdmf.hpp:8-57,
// benchmark.hpp:106-116).
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>

#include "dpa_host.hpp"
#include "json.hpp"

#ifdef DPAH_WITH_HDF5
#include <hdf5.h>
#endif

namespace dpa_hip {

// ------------------------------------------------------------ arguments ----

static std::size_t as_size_string(const json& v) {
  // "Due to limitations of JSON, 64 bit integers are stored as strings."
  const std::string& s = v.as_string();
  size_t pos = 0;
  const unsigned long long x = std::stoull(s, &pos);
  if (pos != s.size()) throw error(DPAH_ERR_INVALID, "not an integer string: " + s);
  return static_cast<std::size_t>(x);
}

static vector3 as_vector3(const json& v) {
  return {static_cast<float>(v[0].as_number()), static_cast<float>(v[1].as_number()),
          static_cast<float>(v[2].as_number())};
}

arguments argument_parser::parse_text(const std::string& text) {
  const json j = json::parse(text);
  arguments a;
  a.input_dataset_filepath = j["input_dataset_filepath"].as_string();
  a.input_dataset_name = j["input_dataset_name"].as_string();
  a.input_dataset_spacing_name = j["input_dataset_spacing_name"].as_string();
  a.particle_advector_load_balancer = j["particle_advector_load_balancer"].as_string();
  a.particle_advector_integrator = j["particle_advector_integrator"].as_string();
  a.particle_advector_step_size = static_cast<float>(j["particle_advector_step_size"].as_number());
  a.particle_advector_gather_particles = j["particle_advector_gather_particles"].as_bool();
  a.particle_advector_record = j["particle_advector_record"].as_bool();
  a.estimate_ftle = j["estimate_ftle"].as_bool();
  a.output_dataset_filepath = j["output_dataset_filepath"].as_string();
  a.seed_generation_iterations = as_size_string(j["seed_generation_iterations"]);
  a.particle_advector_particles_per_round =
      as_size_string(j["particle_advector_particles_per_round"]);
  if (j.contains("seed_generation_stride"))
    a.seed_generation_stride = as_vector3(j["seed_generation_stride"]);
  if (j.contains("seed_generation_count"))
    a.seed_generation_count = as_size_string(j["seed_generation_count"]);
  if (j.contains("seed_generation_range"))
    a.seed_generation_range = std::array<std::size_t, 2>{as_size_string(j["seed_generation_range"][0]),
                                                  as_size_string(j["seed_generation_range"][1])};
  if (j.contains("seed_generation_boundaries")) {
    const json& b = j["seed_generation_boundaries"];
    a.seed_generation_boundaries = aabb3{as_vector3(b["minimum"]), as_vector3(b["maximum"])};
  }
  if (j.contains("seed_generation_random_seed"))
    a.seed_generation_random_seed = as_size_string(j["seed_generation_random_seed"]);
  return a;
}

arguments argument_parser::parse(const std::string& filepath) {
  std::ifstream f(filepath);
  if (!f) throw error(DPAH_ERR_INVALID, "cannot open configuration " + filepath);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_text(ss.str());
}

static std::string fmt_vec(const vector3& v) {
  std::ostringstream s;
  s.precision(9);
  s << "[" << v[0] << ", " << v[1] << ", " << v[2] << "]";
  return s.str();
}

std::string to_json(const arguments& a) {
  std::ostringstream s;
  s.precision(9);
  s << "{\"input_dataset_filepath\": \"" << a.input_dataset_filepath << "\""
    << ", \"input_dataset_name\": \"" << a.input_dataset_name << "\""
    << ", \"input_dataset_spacing_name\": \"" << a.input_dataset_spacing_name << "\""
    << ", \"seed_generation_iterations\": " << a.seed_generation_iterations
    << ", \"particle_advector_particles_per_round\": " << a.particle_advector_particles_per_round
    << ", \"particle_advector_load_balancer\": \"" << a.particle_advector_load_balancer << "\""
    << ", \"particle_advector_integrator\": \"" << a.particle_advector_integrator << "\""
    << ", \"particle_advector_step_size\": " << a.particle_advector_step_size
    << ", \"particle_advector_gather_particles\": " << (a.particle_advector_gather_particles ? "true" : "false")
    << ", \"particle_advector_record\": " << (a.particle_advector_record ? "true" : "false")
    << ", \"estimate_ftle\": " << (a.estimate_ftle ? "true" : "false")
    << ", \"output_dataset_filepath\": \"" << a.output_dataset_filepath << "\"";
  if (a.seed_generation_stride) s << ", \"seed_generation_stride\": " << fmt_vec(*a.seed_generation_stride);
  if (a.seed_generation_count) s << ", \"seed_generation_count\": " << *a.seed_generation_count;
  if (a.seed_generation_range)
    s << ", \"seed_generation_range\": [" << (*a.seed_generation_range)[0] << ", "
      << (*a.seed_generation_range)[1] << "]";
  if (a.seed_generation_boundaries)
    s << ", \"seed_generation_boundaries\": {\"minimum\": "
      << fmt_vec(a.seed_generation_boundaries->minimum)
      << ", \"maximum\": " << fmt_vec(a.seed_generation_boundaries->maximum) << "}";
  s << "}";
  return s.str();
}

// ----------------------------------------------------------------- XDMF ----
// xdmf.hpp:8-57: the same document structure, written by our own templates.

static const char* kXdmfHeader =
    "<?xml version =\"1.0\" ?>\n<!DOCTYPE xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n"
    "  <Domain>\n";
static const char* kXdmfFooter = "\n  </Domain>\n</Xdmf>\n";

static std::string xdmf_polyline(const std::string& file, std::size_t round, std::size_t vertex_floats,
                                 std::size_t color_count, std::size_t index_count, bool u64) {
  std::ostringstream s;
  s << "\n    <Grid Name=\"grid_" << round << "\">\n\n"
    << "      <Topology TopologyType=\"Polyline\" NodesPerElement=\"2\" NumberOfElements=\""
    << index_count / 2 << "\">\n"
    << "        <DataItem Dimensions=\"" << index_count << "\" NumberType=\"UInt\" Precision=\""
    << (u64 ? 8 : 4) << "\" Format=\"HDF\">\n"
    << "          " << file << ":/indices_" << round << "\n        </DataItem>\n      </Topology>\n\n"
    << "      <Geometry GeometryType=\"XYZ\">\n"
    << "        <DataItem Dimensions=\"" << vertex_floats
    << "\" NumberType=\"Float\" Precision=\"4\" Format=\"HDF\">\n"
    << "          " << file << ":/vertices_" << round << "\n        </DataItem>\n      </Geometry>\n\n"
    << "      <Attribute Name=\"Colors\" AttributeType=\"Scalar\" Center=\"Node\">\n"
    << "        <DataItem Dimensions=\"" << color_count
    << "\" NumberType=\"Float\" Precision=\"4\" Format=\"HDF\">\n"
    << "          " << file << ":/colors_" << round << "\n        </DataItem>\n      </Attribute>\n\n"
    << "    </Grid>\n";
  return s.str();
}

static std::string xdmf_volume(const std::string& file, const regular_scalar_field_3d& f) {
  std::ostringstream s;
  const std::string sz = std::to_string(f.shape[0]) + " " + std::to_string(f.shape[1]) + " " +
                         std::to_string(f.shape[2]);
  auto v3 = [](const vector3& v) {
    return std::to_string(v[0]) + " " + std::to_string(v[1]) + " " + std::to_string(v[2]);
  };
  s << "\n    <Grid Name=\"Grid\" GridType=\"Uniform\">\n"
    << "      <Topology TopologyType=\"3DCoRectMesh\" NumberOfElements=\"" << sz << "\"/>\n"
    << "      <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n"
    << "        <DataItem Dimensions=\"3\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n"
    << "          " << v3(f.offset) << "\n        </DataItem>\n"
    << "        <DataItem Dimensions=\"3\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n"
    << "          " << v3(f.spacing) << "\n        </DataItem>\n      </Geometry>\n"
    << "      <Attribute Name=\"Volume\" AttributeType=\"Scalar\" Center=\"Node\">\n"
    << "        <DataItem Dimensions=\"" << sz
    << "\" NumberType=\"Float\" Precision=\"4\" Format=\"HDF\">\n"
    << "          " << file << ":/data\n        </DataItem>\n      </Attribute>\n    </Grid>\n";
  return s.str();
}

// ----------------------------------------------------------------- HDF5 ----

bool hdf5_available() {
#ifdef DPAH_WITH_HDF5
  return true;
#else
  return false;
#endif
}

#ifdef DPAH_WITH_HDF5
namespace {
struct h5 {   // closes an HDF5 handle
  hid_t id;
  herr_t (*close)(hid_t);
  ~h5() { if (id >= 0) close(id); }
};
void h5check(bool ok, const std::string& what) {
  if (!ok) throw error(DPAH_ERR_INVALID, "HDF5: " + what);
}
void write_dataset(hid_t file, const std::string& name, hid_t type, int rank,
                   const hsize_t* dims, const void* data) {
  h5 space{H5Screate_simple(rank, dims, nullptr), H5Sclose};
  h5check(space.id >= 0, "dataspace " + name);
  h5 ds{H5Dcreate2(file, name.c_str(), type, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
        H5Dclose};
  h5check(ds.id >= 0, "create " + name);
  h5check(H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0, "write " + name);
}
}  // namespace
#endif

regular_grid_loader::regular_grid_loader(std::string f, std::string d, std::string s)
    : filepath(std::move(f)), dataset(std::move(d)), spacing(std::move(s)) {}

std::array<std::size_t, 3> regular_grid_loader::load_dimensions() const {
#ifdef DPAH_WITH_HDF5
  h5 file{H5Fopen(filepath.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose};
  h5check(file.id >= 0, "open " + filepath);
  h5 ds{H5Dopen2(file.id, dataset.c_str(), H5P_DEFAULT), H5Dclose};
  h5check(ds.id >= 0, "open dataset " + dataset);
  h5 space{H5Dget_space(ds.id), H5Sclose};
  hsize_t dims[4] = {0, 0, 0, 0};
  h5check(H5Sget_simple_extent_ndims(space.id) == 4, "dataset must be [d0][d1][d2][3]");
  H5Sget_simple_extent_dims(space.id, dims, nullptr);
  h5check(dims[3] == 3, "last dimension must be 3");
  return {dims[0], dims[1], dims[2]};
#else
  throw error(DPAH_ERR_UNSUPPORTED, "built without HDF5");
#endif
}

// regular_grid_loader.cpp:65-88 on one rank: the whole grid, offset 0,
// size = shape * spacing, spacing from a 3-float attribute on the file.
regular_vector_field_3d regular_grid_loader::load_vector_field() const {
#ifdef DPAH_WITH_HDF5
  regular_vector_field_3d f;
  f.shape = load_dimensions();
  f.data.resize(f.shape[0] * f.shape[1] * f.shape[2] * 3);
  h5 file{H5Fopen(filepath.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose};
  h5 ds{H5Dopen2(file.id, dataset.c_str(), H5P_DEFAULT), H5Dclose};
  h5check(H5Dread(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, f.data.data()) >= 0,
          "read " + dataset);
  h5 attr{H5Aopen(file.id, spacing.c_str(), H5P_DEFAULT), H5Aclose};
  h5check(attr.id >= 0, "open attribute " + spacing);
  h5check(H5Aread(attr.id, H5T_NATIVE_FLOAT, f.spacing.data()) >= 0, "read " + spacing);
  for (int i = 0; i < 3; ++i) {
    f.offset[i] = 0.0f * f.spacing[i];
    f.size[i] = static_cast<float>(f.shape[i]) * f.spacing[i];
  }
  return f;
#else
  throw error(DPAH_ERR_UNSUPPORTED, "built without HDF5");
#endif
}

void write_vector_field_h5(const std::string& path, const std::string& dataset,
                           const std::string& spacing_attribute,
                           const regular_vector_field_3d& f) {
#ifdef DPAH_WITH_HDF5
  h5 file{H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
  h5check(file.id >= 0, "create " + path);
  const hsize_t dims[4] = {f.shape[0], f.shape[1], f.shape[2], 3};
  write_dataset(file.id, dataset, H5T_NATIVE_FLOAT, 4, dims, f.data.data());
  const hsize_t three = 3;
  h5 space{H5Screate_simple(1, &three, nullptr), H5Sclose};
  h5 attr{H5Acreate2(file.id, spacing_attribute.c_str(), H5T_NATIVE_FLOAT, space.id, H5P_DEFAULT,
                     H5P_DEFAULT),
          H5Aclose};
  h5check(attr.id >= 0, "create attribute " + spacing_attribute);
  h5check(H5Awrite(attr.id, H5T_NATIVE_FLOAT, f.spacing.data()) >= 0, "write attribute");
#else
  (void)path; (void)dataset; (void)spacing_attribute; (void)f;
  throw error(DPAH_ERR_UNSUPPORTED, "built without HDF5");
#endif
}

// integral_curve_saver.cpp:19-24: <path without extension>.rank_<r>.h5
integral_curve_saver::integral_curve_saver(const std::string& path, int rank)
    : filepath(std::filesystem::path(path).replace_extension(".rank_" + std::to_string(rank) +
                                                             ".h5").string()) {}

void integral_curve_saver::save(const integral_curves& curves) const {
#ifdef DPAH_WITH_HDF5
  std::string bodies;
  {
    h5 file{H5Fcreate(filepath.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
    h5check(file.id >= 0, "create " + filepath);
    const std::string name = std::filesystem::path(filepath).filename().string();
    for (std::size_t r = 0; r < curves.size(); ++r) {
      const integral_curve& c = curves[r];
      if (c.vertices.empty()) continue;                 // empty rounds are skipped (201-202)
      const hsize_t vf = 3 * c.vertices.size();
      write_dataset(file.id, "vertices_" + std::to_string(r), H5T_NATIVE_FLOAT, 1, &vf,
                    c.vertices.data()->data());
      const auto& col = std::get<std::vector<scalar>>(c.colors);
      const hsize_t cf = col.size();
      write_dataset(file.id, "colors_" + std::to_string(r), H5T_NATIVE_FLOAT, 1, &cf, col.data());
      const bool u64 = std::holds_alternative<std::vector<std::uint64_t>>(c.indices);
      hsize_t ic;
      if (u64) {
        const auto& ix = std::get<std::vector<std::uint64_t>>(c.indices);
        ic = ix.size();
        write_dataset(file.id, "indices_" + std::to_string(r), H5T_NATIVE_UINT64, 1, &ic, ix.data());
      } else {
        const auto& ix = std::get<std::vector<std::uint32_t>>(c.indices);
        ic = ix.size();
        write_dataset(file.id, "indices_" + std::to_string(r), H5T_NATIVE_UINT32, 1, &ic, ix.data());
      }
      bodies += xdmf_polyline(name, r, vf, cf, ic, u64);
    }
  }
  if (bodies.empty()) return;
  std::ofstream x(filepath + ".xdmf");
  x << kXdmfHeader << bodies << kXdmfFooter;
#else
  (void)curves;
  throw error(DPAH_ERR_UNSUPPORTED, "built without HDF5");
#endif
}

// regular_grid_saver.cpp:14-46: <path without extension>.rank_<r>_grid.h5
regular_grid_saver::regular_grid_saver(const std::string& path, int rank)
    : filepath(std::filesystem::path(path).replace_extension(".rank_" + std::to_string(rank) +
                                                             "_grid.h5").string()) {}

void regular_grid_saver::save(const regular_scalar_field_3d& f) const {
#ifdef DPAH_WITH_HDF5
  if (f.data.empty()) return;
  {
    h5 file{H5Fcreate(filepath.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
    h5check(file.id >= 0, "create " + filepath);
    const hsize_t dims[3] = {f.shape[0], f.shape[1], f.shape[2]};
    write_dataset(file.id, "data", H5T_NATIVE_FLOAT, 3, dims, f.data.data());
  }
  std::ofstream x(filepath + ".xdmf");
  x << kXdmfHeader << xdmf_volume(std::filesystem::path(filepath).filename().string(), f)
    << kXdmfFooter;
#else
  (void)f;
  throw error(DPAH_ERR_UNSUPPORTED, "built without HDF5");
#endif
}

// ------------------------------------------------------------------ CSV ----

void session_recorder::set(const std::string& name, double value) {
  for (auto& r : records)
    if (r.first == name) { r.second = value; return; }
  records.emplace_back(name, value);
}

// mpi_session::to_csv (benchmark.hpp:106-116) with one rank and one iteration.
void session_recorder::to_csv(const std::string& filepath, int rank) const {
  std::ofstream s(filepath);
  s.precision(std::numeric_limits<double>::max_digits10);
  s << "rank,name,iteration 0,\n";
  for (auto& r : records) s << rank << "," << r.first << "," << r.second << ",\n";
}

}  // namespace dpa_hip
