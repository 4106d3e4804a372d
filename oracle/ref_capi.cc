// extern "C" entry points over the reference deepmd-kit CPU library.
//
// This file is compiled together with the reference sources under
// source/lib/src into oracle/_ref/libdeepmd_ref.so (recipe: oracle/Makefile).
// It only adapts calling conventions (C++ templates, std::vector, InputNlist)
// to plain pointers so that the test-suite and the CPU baseline can drive the
// reference functions through ctypes.  Nothing here re-implements an
// algorithm: every call forwards to the reference function named beside it.
//
// Test infrastructure only: the product library never links this.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "coord.h"
#include "fmt_nlist.h"
#include "neighbor_list.h"
#include "prod_env_mat.h"
#include "prod_force.h"
#include "prod_force_grad.h"
#include "prod_virial.h"
#include "prod_virial_grad.h"
#include "region.h"
#include "tabulate.h"

namespace {
thread_local std::string g_err;
template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// deepmd::init_region_cpu (region.cc:104-108) + normalize_coord_cpu
// (coord.cc:14-28)
int ref_normalize_coord(double* coord, int natom, const double* box) {
  return guard([&] {
    deepmd::Region<double> region;
    deepmd::init_region_cpu(region, box);
    deepmd::normalize_coord_cpu(coord, natom, region);
    return 0;
  });
}

// deepmd::compute_cell_info (coord.cc:69-108), 23 ints.
int ref_compute_cell_info(int* cell_info, float rcut, const double* box) {
  return guard([&] {
    deepmd::Region<double> region;
    deepmd::init_region_cpu(region, box);
    deepmd::compute_cell_info(cell_info, rcut, region);
    return 0;
  });
}

// deepmd::copy_coord_cpu (coord.cc:31-66).  Returns 0, or 1 when
// nall > mem_nall (nall is still written).
int ref_copy_coord(double* out_c, int* out_t, int* mapping, int* nall,
                   const double* in_c, const int* in_t, int nloc,
                   int mem_nall, float rcut, const double* box) {
  return guard([&] {
    deepmd::Region<double> region;
    deepmd::init_region_cpu(region, box);
    return deepmd::copy_coord_cpu(out_c, out_t, mapping, nall, in_c, in_t,
                                  nloc, mem_nall, rcut, region);
  });
}

// deepmd::build_nlist_cpu (neighbor_list.cc:876-928).  firstneigh rows are
// the contiguous slices jlist + ii * mem_size.
int ref_build_nlist(int* ilist, int* numneigh, int* jlist, int* max_list_size,
                    const double* c_cpy, int nloc, int nall, int mem_size,
                    float rcut, int nframes, const int* type) {
  return guard([&] {
    const int nrows = nframes * nloc;
    std::vector<int*> firstneigh(nrows);
    for (int ii = 0; ii < nrows; ++ii) {
      firstneigh[ii] = jlist + (size_t)ii * mem_size;
    }
    deepmd::InputNlist nl(nrows, ilist, numneigh, firstneigh.data());
    return deepmd::build_nlist_cpu(nl, max_list_size, c_cpy, nloc, nall,
                                   mem_size, rcut, nframes, type);
  });
}

// deepmd::format_nlist_cpu (fmt_nlist.cc:146-177), one frame.
int ref_format_nlist(int* nlist, int inum, const int* ilist,
                     const int* numneigh, const int* jlist, int jstride,
                     const double* coord, const int* type, int nloc, int nall,
                     float rcut, const int* sel, int ntypes) {
  return guard([&] {
    std::vector<int*> firstneigh(inum);
    for (int ii = 0; ii < inum; ++ii) {
      firstneigh[ii] = const_cast<int*>(jlist) + (size_t)ii * jstride;
    }
    deepmd::InputNlist nl(inum, const_cast<int*>(ilist),
                          const_cast<int*>(numneigh), firstneigh.data());
    std::vector<int> sec(ntypes + 1, 0);
    for (int tt = 0; tt < ntypes; ++tt) sec[tt + 1] = sec[tt] + sel[tt];
    deepmd::format_nlist_cpu(nlist, nl, coord, type, nloc, nall, rcut, sec);
    return 0;
  });
}

// deepmd::prod_env_mat_a_cpu (prod_env_mat.cc:15-132).
int ref_prod_env_mat_a(double* em, double* em_deriv, double* rij, int* nlist,
                       const double* coord, const int* type, int inum,
                       const int* ilist, const int* numneigh, const int* jlist,
                       int jstride, int max_nbor_size, const double* avg,
                       const double* std, int nloc, int nall, int nframes,
                       float rcut, float rcut_smth, const int* sel,
                       int ntypes) {
  return guard([&] {
    std::vector<int*> firstneigh(inum);
    for (int ii = 0; ii < inum; ++ii) {
      firstneigh[ii] = const_cast<int*>(jlist) + (size_t)ii * jstride;
    }
    deepmd::InputNlist nl(inum, const_cast<int*>(ilist),
                          const_cast<int*>(numneigh), firstneigh.data());
    std::vector<int> sec(ntypes + 1, 0);
    for (int tt = 0; tt < ntypes; ++tt) sec[tt + 1] = sec[tt] + sel[tt];
    deepmd::prod_env_mat_a_cpu(em, em_deriv, rij, nlist, coord, type, nl,
                               max_nbor_size, avg, std, nloc, nall, nframes,
                               rcut, rcut_smth, sec);
    return 0;
  });
}

#define REF_TAB(SUF, T)                                                       \
  int ref_tabulate_fusion_se_a_##SUF(T* out, const T* table,                  \
                                     const T* table_info, const T* em_x,      \
                                     const T* em, int nloc, int nnei,         \
                                     int last_layer_size, int is_sorted,      \
                                     int ndescrpt) {                          \
    return guard([&] {                                                        \
      deepmd::tabulate_fusion_se_a_cpu<T>(out, table, table_info, em_x, em,   \
                                          nullptr, nloc, nnei,                \
                                          last_layer_size, is_sorted != 0,    \
                                          ndescrpt);                          \
      return 0;                                                               \
    });                                                                       \
  }                                                                           \
  int ref_tabulate_fusion_se_a_grad_##SUF(                                    \
      T* dy_dem_x, T* dy_dem, const T* table, const T* table_info,            \
      const T* em_x, const T* em, const T* dy, int nloc, int nnei,            \
      int last_layer_size, int is_sorted, int ndescrpt) {                     \
    return guard([&] {                                                        \
      deepmd::tabulate_fusion_se_a_grad_cpu<T>(                               \
          dy_dem_x, dy_dem, nullptr, table, table_info, em_x, em, nullptr,    \
          dy, nloc, nnei, last_layer_size, is_sorted != 0, ndescrpt);         \
      return 0;                                                               \
    });                                                                       \
  }                                                                           \
  int ref_tabulate_fusion_se_a_grad_grad_##SUF(                               \
      T* dz_dy, const T* table, const T* table_info, const T* em_x,           \
      const T* em, const T* dz_dy_dem_x, const T* dz_dy_dem, int nloc,        \
      int nnei, int last_layer_size, int is_sorted, int ndescrpt) {           \
    return guard([&] {                                                        \
      deepmd::tabulate_fusion_se_a_grad_grad_cpu<T>(                          \
          dz_dy, table, table_info, em_x, em, nullptr, dz_dy_dem_x,           \
          dz_dy_dem, nullptr, nloc, nnei, last_layer_size, is_sorted != 0,    \
          ndescrpt);                                                          \
      return 0;                                                               \
    });                                                                       \
  }

// deepmd::tabulate_fusion_se_a{,_grad,_grad_grad}_cpu (tabulate.cc:451-566)
REF_TAB(f64, double)
REF_TAB(f32, float)

// deepmd::prod_force_a_cpu (prod_force.cc:76-87)
int ref_prod_force_a(double* force, const double* net_deriv,
                     const double* env_deriv, const int* nlist, int nloc,
                     int nall, int nnei, int nframes) {
  return guard([&] {
    deepmd::prod_force_a_cpu(force, net_deriv, env_deriv, nlist, nloc, nall,
                             nnei, nframes);
    return 0;
  });
}

// deepmd::prod_virial_a_cpu (prod_virial.cc:23-69), one frame.
int ref_prod_virial_a(double* virial, double* atom_virial,
                      const double* net_deriv, const double* env_deriv,
                      const double* rij, const int* nlist, int nloc, int nall,
                      int nnei) {
  return guard([&] {
    deepmd::prod_virial_a_cpu(virial, atom_virial, net_deriv, env_deriv, rij,
                              nlist, nloc, nall, nnei);
    return 0;
  });
}

// deepmd::prod_force_grad_a_cpu (prod_force_grad.cc:23-75)
int ref_prod_force_grad_a(double* grad_net, const double* grad,
                          const double* env_deriv, const int* nlist, int nloc,
                          int nnei, int nframes) {
  return guard([&] {
    deepmd::prod_force_grad_a_cpu(grad_net, grad, env_deriv, nlist, nloc,
                                  nnei, nframes);
    return 0;
  });
}

// deepmd::prod_virial_grad_a_cpu (prod_virial_grad.cc:22-66)
int ref_prod_virial_grad_a(double* grad_net, const double* grad,
                           const double* env_deriv, const double* rij,
                           const int* nlist, int nloc, int nnei) {
  return guard([&] {
    deepmd::prod_virial_grad_a_cpu(grad_net, grad, env_deriv, rij, nlist,
                                   nloc, nnei);
    return 0;
  });
}

}  // extern "C"
