// dpmi model files (layout in model_file.cpp) and the native model build.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace dpmi {

struct SeaModel;

struct ModelFile {
  struct Net {
    std::vector<std::vector<double>> w, b, idt;  // per layer; idt empty when absent
  };
  int ntypes = 0, type_one_side = 0, axis = 0;
  double rcut = 0, rcut_smth = 0;
  std::vector<int> sel, embed_width, fit_width, fit_act, fit_resnet;
  std::vector<double> davg, dstd, bias_atom_e;
  std::vector<Net> embed, fit;  // per net / per type
  double min_nbor_dist = 0, extrapolate = 5, stride0 = 0.01, stride1 = 0.1, check_frequency = -1;
  std::vector<std::string> type_map;
};

ModelFile read_model_file(const std::string& path);
// (lower, upper) of every embedding net's table
std::vector<std::pair<double, double>> model_table_ranges(const ModelFile& m);
// compressed tables built on the device, then the engine model
SeaModel* build_model(const ModelFile& m, hipStream_t st);

}  // namespace dpmi
