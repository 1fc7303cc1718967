// main.cpp -- the dpa_hip command line (source/main.cpp:5-8 for one GPU).
//
//   dpa_hip <config.json>                    run one rank's pipeline on GPU 0
//   dpa_hip --check-config <config.json>     parse the configuration, print it as JSON
//   dpa_hip --make-abc-field <out.h5> <n0> <n1> <n2> [dataset] [spacing_attribute]
//                                            write an ABC-flow field in dpa's HDF5 layout
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "dpa_host.hpp"

namespace {

int make_abc(int argc, char** argv) {
  if (argc < 6) {
    std::cerr << "usage: dpa_hip --make-abc-field <out.h5> <n0> <n1> <n2> [dataset] [spacing]\n";
    return 2;
  }
  dpa_hip::regular_vector_field_3d f;
  for (int i = 0; i < 3; ++i) f.shape[i] = std::strtoull(argv[3 + i], nullptr, 10);
  const std::string dataset = argc > 6 ? argv[6] : "Data3";
  const std::string spacing = argc > 7 ? argv[7] : "spacing";
  const double A = std::sqrt(3.0), B = std::sqrt(2.0), C = 1.0, two_pi = 2.0 * M_PI;
  for (int i = 0; i < 3; ++i) f.spacing[i] = static_cast<float>(two_pi / double(f.shape[i] - 1));
  f.data.resize(f.shape[0] * f.shape[1] * f.shape[2] * 3);
  size_t k = 0;
  for (size_t i0 = 0; i0 < f.shape[0]; ++i0)        // axis 0 = z
    for (size_t i1 = 0; i1 < f.shape[1]; ++i1)      // axis 1 = y
      for (size_t i2 = 0; i2 < f.shape[2]; ++i2) {  // axis 2 = x
        const double z = i0 * double(f.spacing[0]), y = i1 * double(f.spacing[1]),
                     x = i2 * double(f.spacing[2]);
        f.data[k++] = static_cast<float>(A * std::sin(z) + C * std::cos(y));
        f.data[k++] = static_cast<float>(B * std::sin(x) + A * std::cos(z));
        f.data[k++] = static_cast<float>(C * std::sin(y) + B * std::cos(x));
      }
  dpa_hip::write_vector_field_h5(argv[2], dataset, spacing, f);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::string(argv[1]) == "--check-config") {
      if (argc < 3) return 2;
      std::cout << dpa_hip::to_json(dpa_hip::argument_parser::parse(argv[2])) << "\n";
      return 0;
    }
    if (argc >= 2 && std::string(argv[1]) == "--make-abc-field") return make_abc(argc, argv);
    return dpa_hip::pipeline::run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "dpa_hip: " << e.what() << "\n";
    return 1;
  }
}
