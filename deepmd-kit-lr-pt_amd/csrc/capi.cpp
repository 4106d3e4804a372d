// extern "C" entry points of include/dpmi.h.
//
// Each kernel-level entry point keeps the argument meaning, output layout and
// error behaviour of the reference function it replaces (cited in
// include/dpmi.h); errors are returned as dpmi_status with a message in
// dpmi_last_error() instead of thrown.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "dpmi_common.h"
#include "engine.h"
#include "kernels.h"
#include "model_file.h"

namespace {

constexpr int kOk = DPMI_OK;
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const dpmi::Unsupported& e) {
    return fail(DPMI_ERR_UNSUPPORTED, e.what());
  } catch (const dpmi::HipError& e) {
    return fail(DPMI_ERR_HIP, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(DPMI_ERR_ARGUMENT, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(DPMI_ERR_MEMORY, e.what());
  } catch (const std::exception& e) {
    return fail(DPMI_ERR_HIP, e.what());
  }
}

// Grow-only device scratch per purpose for the kernel-level entry points
// (the reference takes such workspaces from its caller).
struct Scratch {
  std::mutex mu;
  std::map<std::string, std::pair<void*, size_t>> bufs;
  void* get(const std::string& key, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu);
    auto& b = bufs[key];
    if (b.second < bytes) {
      if (b.first) (void)hipFree(b.first);
      b.first = nullptr;
      DPMI_CHECK(hipMalloc(&b.first, std::max<size_t>(bytes, 256)));
      b.second = bytes;
    }
    return b.first;
  }
};
Scratch& scratch() {
  static Scratch* s = new Scratch();
  return *s;
}

hipStream_t S(void* st) { return static_cast<hipStream_t>(st); }

int flag_to_status(int flag) {
  if (flag & dpmi::kFlagNlistCapacity) return fail(DPMI_ERR_NLIST_CAPACITY, "neighbour row capacity exceeded");
  if (flag & dpmi::kFlagGhostCapacity) return fail(DPMI_ERR_MEMORY, "nall exceeds mem_nall");
  if (flag & dpmi::kFlagIndexRange) return fail(DPMI_ERR_RANGE, "extended index >= 2^24 in the neighbour sort key");
  if (flag & dpmi::kFlagTypeRange) return fail(DPMI_ERR_RANGE, "atom type >= ntypes in the neighbour sort key");
  if (flag & dpmi::kFlagHalfRange)
    return fail(DPMI_ERR_RANGE, "a mixed-precision matrix-core operand exceeds the FP16 range");
  if (flag & dpmi::kFlagTypeMismatch)
    return fail(DPMI_ERR_ARGUMENT, "the local atom types differ from the types the engine was created for");
  if (flag & dpmi::kFlagGridChanged)
    return fail(DPMI_ERR_ARGUMENT,
                "a replayed box needs another cell grid than the captured one (the step used the old box): capture "
                "again");
  return kOk;
}

int check_tab(int nd, int M) {
  if (!(nd == 4 || nd == 9 || nd == 16 || nd == 25))
    return fail(DPMI_ERR_UNSUPPORTED,
                "The environment basis dimension must be 4, 9, 16, or 25, got " + std::to_string(nd));
  if (M <= 0 || M > 1024)
    return fail(DPMI_ERR_UNSUPPORTED,
                "last_layer_size must be between 1 and 1024 for GPU tabulation, but got " + std::to_string(M));
  return kOk;
}

}  // namespace

extern "C" {

const char* dpmi_last_error(void) { return g_err.c_str(); }

const char* dpmi_version(void) { return "dpmi 0.1 gfx950 (fp64 engine, hip kernels)"; }

int dpmi_normalize_coord_f64(double* coord, int natom, const double* boxt_host, void* stream) {
  return guarded([&] {
    dpmi::FrameGeom g;
    dpmi::geom_from_box(g, boxt_host, 1.0f);
    dpmi::launch_normalize(coord, natom, g, S(stream));
    return kOk;
  });
}

int dpmi_copy_coord_f64(double* out_c, int* out_t, int* mapping, int* nall_host, const double* in_c, const int* in_t,
                        int nloc, int mem_nall, float rcut, const double* boxt_host, void* stream) {
  return guarded([&] {
    dpmi::FrameGeom g;
    dpmi::geom_from_box(g, boxt_host, rcut);
    return dpmi::copy_coord_run(g, out_c, out_t, mapping, nall_host, in_c, in_t, nloc, mem_nall, stream);
  });
}

}  // extern "C"

namespace dpmi {
// Single-frame ghost copy on the grid of g (cell sizes and ghost layers),
// used by dpmi_copy_coord_f64 and by deepmd::copy_coord_gpu (whose grid comes
// from the caller's cell_info).
int copy_coord_run(FrameGeom g, double* out_c, int* out_t, int* mapping, int* nall_host, const double* in_c,
                   const int* in_t, int nloc, int mem_nall, void* stream) {
  {
    if (nloc < 0 || mem_nall < nloc) throw std::invalid_argument("mem_nall must be >= nloc >= 0");
    hipStream_t st = S(stream);
    g.nloc = nloc;
    g.loc_off = 0;
    g.ext_off = 0;
    g.ext_cap = mem_nall;
    g.cell_off = 0;
    g.xcell_off = 0;
    const int nc = dpmi::geom_local_cells(g), nx = dpmi::geom_ext_cells(g) + 1;
    auto& sc = scratch();
    // [FrameGeom][pad to 16][CellAtom x mem_nall][ints: 5 of nloc, 3 of nc,
    // 5 of nx, nall, flag]
    const size_t rec_off = (sizeof(dpmi::FrameGeom) + 15) / 16 * 16;
    char* base = static_cast<char*>(sc.get("copy_coord", rec_off + sizeof(dpmi::CellAtom) * (size_t)mem_nall +
                                                              sizeof(int) * ((size_t)5 * nloc + 3 * nc + 5 * nx + 8)));
    dpmi::CellAtom* recs = reinterpret_cast<dpmi::CellAtom*>(base + rec_off);
    int* ip = reinterpret_cast<int*>(recs + mem_nall);
    auto take = [&](size_t n) {
      int* p = ip;
      ip += n;
      return p;
    };
    dpmi::CellWork w;
    w.nframes = 1;
    w.ntot = nloc;
    w.max_nloc = nloc;
    w.ncells_tot = nc;
    w.nx_tot = nx;
    w.do_wrap = 0;
    w.coord_in = in_c;
    w.type_in = in_t;
    int* frame_of = take(nloc);
    int* cell_frame = take(nc);
    int* xcell_frame = take(nx);
    w.frame_of = frame_of;
    w.cell_frame = cell_frame;
    w.xcell_frame = xcell_frame;
    w.geom = reinterpret_cast<dpmi::FrameGeom*>(base);
    w.coord_ext = out_c;
    w.type_ext = out_t;
    w.map_ext = mapping;
    w.cell_atoms = recs;
    w.cell_of = take(nloc);
    w.cell_slot = take(nloc);
    w.cell_cnt = take(nc);
    w.cell_start = take(nc);
    w.cell_list = take(nloc);
    w.sorted_loc = take(nloc);
    w.xcnt = take(nx);
    w.xstart = take(nx);
    w.xgcnt = take(nx);
    w.xgstart = take(nx);
    w.nall = take(1);
    w.flag = take(1);
    DPMI_CHECK(hipMemcpyAsync(base, &g, sizeof(g), hipMemcpyHostToDevice, st));
    DPMI_CHECK(hipMemsetAsync(frame_of, 0, sizeof(int) * nloc, st));
    DPMI_CHECK(hipMemsetAsync(cell_frame, 0, sizeof(int) * nc, st));
    DPMI_CHECK(hipMemsetAsync(xcell_frame, 0, sizeof(int) * nx, st));
    DPMI_CHECK(hipMemsetAsync(w.flag, 0, sizeof(int), st));
    dpmi::launch_build_extended(w, st);
    int out[2];
    DPMI_CHECK(hipMemcpyAsync(out, w.nall, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
    DPMI_CHECK(hipStreamSynchronize(st));
    *nall_host = out[0];
    if (out[0] > mem_nall) return fail(DPMI_ERR_MEMORY, "nall exceeds mem_nall");
    return kOk;
  }
}
}  // namespace dpmi

extern "C" {

int dpmi_build_nlist_f64(int* ilist, int* numneigh, int* jlist, int* max_list_size_host, const double* c_cpy,
                         int nloc, int nall, int mem_size, float rcut, int nframes, const int* type, void* stream) {
  return guarded([&] {
    hipStream_t st = S(stream);
    const int rows = nframes * nloc;
    dpmi::launch_brute_nlist(c_cpy, nloc, nall, mem_size, (double)(rcut * rcut), nframes, type, ilist, numneigh,
                             jlist, st);
    std::vector<int> nn(rows);
    if (rows) DPMI_CHECK(hipMemcpyAsync(nn.data(), numneigh, sizeof(int) * rows, hipMemcpyDeviceToHost, st));
    DPMI_CHECK(hipStreamSynchronize(st));
    int mx = 0;
    for (int v : nn) mx = std::max(mx, v);
    *max_list_size_host = mx;
    if (mx > mem_size) return fail(DPMI_ERR_NLIST_CAPACITY, "a neighbour row exceeds mem_size");
    return kOk;
  });
}

int dpmi_prod_env_mat_a_f64(double* em, double* em_deriv, double* rij, int* nlist, const double* coord,
                            const int* type, int inum, const int* ilist, const int* numneigh, const int* jlist,
                            int jstride, const double* avg, const double* std_, int nloc, int nall, int nframes,
                            float rcut, float rcut_smth, const int* sel_host, int ntypes, void* stream) {
  return guarded([&] {
    hipStream_t st = S(stream);
    if (ntypes <= 0 || ntypes > 127) throw std::invalid_argument("ntypes must be in [1, 127]");
    if (inum != nframes * nloc) throw std::invalid_argument("inum must equal nframes * nloc");
    std::vector<int> sec(ntypes + 1, 0);
    for (int t = 0; t < ntypes; ++t) sec[t + 1] = sec[t] + sel_host[t];
    const int nnei = sec[ntypes];
    auto& sc = scratch();
    int* d_sec = static_cast<int*>(sc.get("env_sec", sizeof(int) * (ntypes + 2)));
    int* flag = d_sec + ntypes + 1;
    DPMI_CHECK(hipMemcpyAsync(d_sec, sec.data(), sizeof(int) * (ntypes + 1), hipMemcpyHostToDevice, st));
    DPMI_CHECK(hipMemsetAsync(flag, 0, sizeof(int), st));
    DPMI_CHECK(hipMemsetAsync(nlist, 0xff, sizeof(int) * (size_t)nframes * nloc * nnei, st));
    dpmi::NlistArgs na;
    na.cap = dpmi::nlist_row_capacity(std::max(jstride, 64));
    na.sec = d_sec;
    na.ntypes = ntypes;
    na.nnei = nnei;
    na.rc2f = rcut * rcut;
    na.nlist = nlist;
    dpmi::launch_row_format(inum, ilist, numneigh, jlist, jstride, coord, type, nloc, nall, na, flag, st);
    dpmi::EnvArgs ea;
    ea.coord = coord;
    ea.type = type;
    ea.nlist = nlist;
    ea.strided_nloc = nloc;
    ea.nrows = (long long)nframes * nloc;
    ea.nnei = nnei;
    ea.nall_stride = nall;
    ea.ntypes = ntypes;
    ea.avg = avg;
    ea.std_ = std_;
    ea.rmin = (double)rcut_smth;
    ea.rmax = (double)rcut;
    ea.em = em;
    ea.em_deriv = em_deriv;
    ea.rij = rij;
    dpmi::launch_env_mat_a(ea, st);
    int f = 0;
    DPMI_CHECK(hipMemcpyAsync(&f, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    DPMI_CHECK(hipStreamSynchronize(st));
    return flag_to_status(f);
  });
}

#define DPMI_TAB_ENTRY(SUF, T)                                                                                       \
  int dpmi_tabulate_fusion_se_a_##SUF(T* out, const T* table, const T* info, const T* em_x, const T* em, int nloc, \
                                      int nnei, int M, int is_sorted, int nd, void* stream) {                        \
    return guarded([&] {                                                                                             \
      if (int e = check_tab(nd, M)) return e;                                                                        \
      dpmi::launch_tab_fwd<T>(out, table, info, em_x, em, nloc, nnei, M, is_sorted, nd, S(stream));                  \
      return kOk;                                                                                           \
    });                                                                                                              \
  }                                                                                                                  \
  int dpmi_tabulate_fusion_se_a_grad_##SUF(T* dx, T* dem, const T* table, const T* info, const T* em_x, const T* em, \
                                           const T* dy, int nloc, int nnei, int M, int is_sorted, int nd,           \
                                           void* stream) {                                                           \
    return guarded([&] {                                                                                             \
      if (int e = check_tab(nd, M)) return e;                                                                        \
      dpmi::launch_tab_grad<T>(dx, dem, table, info, em_x, em, dy, nloc, nnei, M, is_sorted, nd, S(stream));        \
      return kOk;                                                                                           \
    });                                                                                                              \
  }                                                                                                                  \
  int dpmi_tabulate_fusion_se_a_grad_grad_##SUF(T* dz, const T* table, const T* info, const T* em_x, const T* em,   \
                                                const T* zx, const T* zem, int nloc, int nnei, int M, int is_sorted, \
                                                int nd, void* stream) {                                              \
    return guarded([&] {                                                                                             \
      if (int e = check_tab(nd, M)) return e;                                                                        \
      dpmi::launch_tab_grad_grad<T>(dz, table, info, em_x, em, zx, zem, nloc, nnei, M, is_sorted, nd, S(stream));   \
      return kOk;                                                                                           \
    });                                                                                                              \
  }

DPMI_TAB_ENTRY(f64, double)
DPMI_TAB_ENTRY(f32, float)

extern "C++" {
namespace {

template <typename T>
int prod_force_impl(T* force, const T* net_deriv, const T* in_deriv, const int* nlist, int nloc, int nall, int nnei,
                    int nframes, void* stream) {
  return guarded([&] {
    dpmi::launch_prod_force_a<T>(force, net_deriv, in_deriv, nlist, nloc, nall, nnei, nframes, S(stream));
    return kOk;
  });
}

template <typename T>
int prod_virial_impl(T* virial, T* atom_virial, const T* net_deriv, const T* env_deriv, const T* rij,
                     const int* nlist, int nloc, int nall, int nnei, void* stream) {
  return guarded([&] {
    hipStream_t st = S(stream);
    double* vpart = static_cast<double*>(scratch().get("virial_part", sizeof(double) * 9 * (size_t)std::max(nloc, 1)));
    dpmi::launch_prod_virial_a<T>(vpart, atom_virial, net_deriv, env_deriv, rij, nlist, nloc, nall, nnei, st);
    if (nloc > 0)
      dpmi::launch_reduce_rows9<T>(virial, vpart, nloc, st);
    else
      DPMI_CHECK(hipMemsetAsync(virial, 0, sizeof(T) * 9, st));
    // the shared partial buffer is reused by the next call: finish here
    DPMI_CHECK(hipStreamSynchronize(st));
    return kOk;
  });
}

template <typename T>
int prod_force_grad_impl(T* grad_net, const T* grad, const T* env_deriv, const int* nlist, int nloc, int nnei,
                         int nframes, void* stream) {
  return guarded([&] {
    dpmi::launch_prod_force_grad_a<T>(grad_net, grad, env_deriv, nlist, nloc, nnei, nframes, S(stream));
    return kOk;
  });
}

template <typename T>
int prod_virial_grad_impl(T* grad_net, const T* grad, const T* env_deriv, const T* rij, const int* nlist, int nloc,
                          int nnei, void* stream) {
  return guarded([&] {
    dpmi::launch_prod_virial_grad_a<T>(grad_net, grad, env_deriv, rij, nlist, nloc, nnei, S(stream));
    return kOk;
  });
}

}  // namespace
}  // extern "C++"

int dpmi_prod_force_a_f64(double* force, const double* net_deriv, const double* in_deriv, const int* nlist, int nloc,
                          int nall, int nnei, int nframes, void* stream) {
  return prod_force_impl(force, net_deriv, in_deriv, nlist, nloc, nall, nnei, nframes, stream);
}
int dpmi_prod_force_a_f32(float* force, const float* net_deriv, const float* in_deriv, const int* nlist, int nloc,
                          int nall, int nnei, int nframes, void* stream) {
  return prod_force_impl(force, net_deriv, in_deriv, nlist, nloc, nall, nnei, nframes, stream);
}
int dpmi_prod_virial_a_f64(double* virial, double* atom_virial, const double* net_deriv, const double* env_deriv,
                           const double* rij, const int* nlist, int nloc, int nall, int nnei, void* stream) {
  return prod_virial_impl(virial, atom_virial, net_deriv, env_deriv, rij, nlist, nloc, nall, nnei, stream);
}
int dpmi_prod_virial_a_f32(float* virial, float* atom_virial, const float* net_deriv, const float* env_deriv,
                           const float* rij, const int* nlist, int nloc, int nall, int nnei, void* stream) {
  return prod_virial_impl(virial, atom_virial, net_deriv, env_deriv, rij, nlist, nloc, nall, nnei, stream);
}
int dpmi_prod_virial_a_f64_noalloc(double* virial, double* atom_virial, const double* net_deriv,
                                   const double* env_deriv, const double* rij, const int* nlist, int nloc, int nall,
                                   int nnei, void* stream) {
  return guarded([&] {
    if (nloc < 0 || nloc > dpmi::kNoAllocVirialRows)
      throw std::invalid_argument("dpmi_prod_virial_a_f64_noalloc: nloc " + std::to_string(nloc) +
                                  " exceeds the fixed scratch of " + std::to_string(dpmi::kNoAllocVirialRows) +
                                  " rows (use dpmi_prod_virial_a_f64)");
    hipStream_t st = S(stream);
    double* vpart = dpmi::noalloc_virial_scratch();
    // every row r < nloc of the scratch is written before it is summed
    dpmi::launch_prod_virial_a<double>(vpart, atom_virial, net_deriv, env_deriv, rij, nlist, nloc, nall, nnei, st);
    if (nloc > 0)
      dpmi::launch_reduce_rows9<double>(virial, vpart, nloc, st);
    else
      DPMI_CHECK(hipMemsetAsync(virial, 0, sizeof(double) * 9, st));
    return kOk;
  });
}
int dpmi_prod_force_grad_a_f64(double* grad_net, const double* grad, const double* env_deriv, const int* nlist,
                               int nloc, int nnei, int nframes, void* stream) {
  return prod_force_grad_impl(grad_net, grad, env_deriv, nlist, nloc, nnei, nframes, stream);
}
int dpmi_prod_force_grad_a_f32(float* grad_net, const float* grad, const float* env_deriv, const int* nlist,
                               int nloc, int nnei, int nframes, void* stream) {
  return prod_force_grad_impl(grad_net, grad, env_deriv, nlist, nloc, nnei, nframes, stream);
}
int dpmi_prod_virial_grad_a_f64(double* grad_net, const double* grad, const double* env_deriv, const double* rij,
                                const int* nlist, int nloc, int nnei, void* stream) {
  return prod_virial_grad_impl(grad_net, grad, env_deriv, rij, nlist, nloc, nnei, stream);
}
int dpmi_prod_virial_grad_a_f32(float* grad_net, const float* grad, const float* env_deriv, const float* rij,
                                const int* nlist, int nloc, int nnei, void* stream) {
  return prod_virial_grad_impl(grad_net, grad, env_deriv, rij, nlist, nloc, nnei, stream);
}

int dpmi_tabulate_build_se_a(double* table_host, int* nspline_host, int nlayers, const int* width_host,
                             const double* const* w_host, const double* const* b_host,
                             const double* const* idt_host, double lower, double upper, double extrapolate,
                             double stride0, double stride1, void* stream) {
  return guarded([&] {
    if (!nspline_host || nlayers < 1 || !width_host || !w_host || !b_host)
      throw std::invalid_argument("null argument");
    const int ns = dpmi::tab_nspline(lower, upper, extrapolate, stride0, stride1);
    if (!table_host) {
      *nspline_host = ns;
      return kOk;
    }
    dpmi::TabNet net;
    net.width.assign(width_host, width_host + nlayers + 1);
    for (int l = 0; l < nlayers; ++l) {
      net.w.push_back(w_host[l]);
      net.b.push_back(b_host[l]);
      net.idt.push_back(idt_host ? idt_host[l] : nullptr);
    }
    std::vector<double> tab;
    int n = 0;
    dpmi::build_table_device(net, lower, upper, extrapolate, stride0, stride1, tab, n, S(stream));
    std::memcpy(table_host, tab.data(), sizeof(double) * tab.size());
    *nspline_host = n;
    return kOk;
  });
}

// ---------------------------------------------------------------- engine --
int dpmi_sea_model_create(dpmi_sea_model** model, const dpmi_sea_model_desc* desc) {
  return guarded([&] {
    if (!model || !desc) throw std::invalid_argument("null argument");
    auto* m = new dpmi_sea_model{nullptr};
    try {
      m->impl = new dpmi::SeaModel(*desc);
    } catch (...) {
      delete m;
      throw;
    }
    *model = m;
    return kOk;
  });
}

int dpmi_sea_model_load(dpmi_sea_model** model, const char* path, void* stream) {
  return guarded([&] {
    if (!model || !path) throw std::invalid_argument("null argument");
    const dpmi::ModelFile mf = dpmi::read_model_file(path);
    auto* m = new dpmi_sea_model{nullptr};
    try {
      m->impl = dpmi::build_model(mf, S(stream));
    } catch (...) {
      delete m;
      throw;
    }
    *model = m;
    return kOk;
  });
}

int dpmi_sea_model_info(const dpmi_sea_model* model, int* ntypes_host, double* rcut_host, int* nnei_host,
                        char* type_map_host, int type_map_len) {
  return guarded([&] {
    if (!model) throw std::invalid_argument("null model");
    const dpmi::SeaModel& m = *model->impl;
    if (ntypes_host) *ntypes_host = m.ntypes;
    if (rcut_host) *rcut_host = (double)m.rcut;
    if (nnei_host) *nnei_host = m.nnei;
    if (type_map_host && type_map_len > 0) {
      std::string s;
      for (size_t k = 0; k < m.type_map.size(); ++k) s += (k ? " " : "") + m.type_map[k];
      std::strncpy(type_map_host, s.c_str(), type_map_len - 1);
      type_map_host[type_map_len - 1] = 0;
    }
    return kOk;
  });
}

void dpmi_sea_model_destroy(dpmi_sea_model* model) {
  if (!model) return;
  delete model->impl;
  delete model;
}

int dpmi_sea_engine_create(dpmi_sea_engine** engine, const dpmi_sea_model* model, int nframes,
                           const int* natoms_host, const int* atype_host, int precision, double mem_nall_per_atom) {
  return guarded([&] {
    if (!engine || !model || nframes <= 0) throw std::invalid_argument("invalid engine arguments");
    if (precision != DPMI_FP64 && precision != DPMI_MIXED) throw std::invalid_argument("unknown precision");
    auto* e = new dpmi_sea_engine{nullptr};
    try {
      e->impl = new dpmi::SeaEngine(model->impl, nframes, natoms_host, atype_host, precision, mem_nall_per_atom);
    } catch (...) {
      delete e;
      throw;
    }
    *engine = e;
    return kOk;
  });
}

void dpmi_sea_engine_destroy(dpmi_sea_engine* engine) {
  if (!engine) return;
  delete engine->impl;
  delete engine;
}

int dpmi_sea_engine_compute(dpmi_sea_engine* engine, const double* coord, const double* box_host, double* energy,
                            double* force, double* virial, double* atom_energy, double* atom_virial, void* stream) {
  return guarded([&] {
    if (!engine || !coord || !box_host || !energy || !force || !virial) throw std::invalid_argument("null argument");
    engine->impl->compute(coord, box_host, energy, force, virial, atom_energy, atom_virial, S(stream));
    return kOk;
  });
}

int dpmi_sea_engine_compute_lower(dpmi_sea_engine* engine, const double* coord_ext, const int* atype_ext, int nall,
                                  int inum, const int* ilist, const int* numneigh, const int* jlist, int jstride,
                                  double* energy, double* force_ext, double* virial, double* atom_energy,
                                  double* atom_virial_ext, void* stream) {
  return guarded([&] {
    if (!engine || !coord_ext || !atype_ext || !ilist || !numneigh || (!jlist && jstride > 0) || !energy ||
        !force_ext || !virial)
      throw std::invalid_argument("null argument");
    engine->impl->compute_lower(coord_ext, atype_ext, nall, inum, ilist, numneigh, jlist, jstride, energy, force_ext,
                                virial, atom_energy, atom_virial_ext, S(stream));
    return kOk;
  });
}

int dpmi_sea_engine_set_growth(dpmi_sea_engine* engine, int enable) {
  return guarded([&] {
    if (!engine) throw std::invalid_argument("null engine");
    engine->impl->grow = enable != 0;
    return kOk;
  });
}

int dpmi_sea_set_net_chunk_bytes(long long bytes) {
  return guarded([&] {
    if (bytes <= 0) throw std::invalid_argument("chunk bytes must be positive");
    dpmi::SeaEngine::net_chunk_bytes = bytes;
    return kOk;
  });
}

int dpmi_sea_engine_graph_capture(dpmi_sea_engine* engine, const double* coord, const double* box_host,
                                  double* energy, double* force, double* virial, double* atom_energy, void* stream) {
  return guarded([&] {
    if (!engine || !coord || !box_host || !energy || !force || !virial) throw std::invalid_argument("null argument");
    engine->impl->capture(coord, box_host, energy, force, virial, atom_energy, S(stream));
    return kOk;
  });
}

int dpmi_sea_engine_graph_capture_device_box(dpmi_sea_engine* engine, const double* coord, const double* box_dev,
                                             double* energy, double* force, double* virial, double* atom_energy,
                                             void* stream) {
  return guarded([&] {
    if (!engine || !coord || !box_dev || !energy || !force || !virial) throw std::invalid_argument("null argument");
    engine->impl->capture_device_box(coord, box_dev, energy, force, virial, atom_energy, S(stream));
    return kOk;
  });
}

int dpmi_sea_engine_graph_launch(dpmi_sea_engine* engine, void* stream) {
  return guarded([&] {
    if (!engine) throw std::invalid_argument("null engine");
    engine->impl->launch_graph(S(stream));
    return kOk;
  });
}

int dpmi_sea_engine_autotune(dpmi_sea_engine* engine, const double* coord, const double* box_host, int reps,
                             void* stream) {
  return guarded([&] {
    if (!engine || !coord || !box_host) throw std::invalid_argument("null argument");
    engine->impl->autotune(coord, box_host, reps, S(stream));
    return kOk;
  });
}

int dpmi_sea_engine_tuning(dpmi_sea_engine* engine, int* knobs_host) {
  return guarded([&] {
    if (!engine || !knobs_host) throw std::invalid_argument("null argument");
    const dpmi::Tune& t = engine->impl->tune;
    knobs_host[0] = t.env_tpb;
    knobs_host[1] = t.fwd_waves;
    knobs_host[2] = t.bwd_waves;
    knobs_host[3] = t.force_rows;
    knobs_host[4] = t.embed_mfma;
    knobs_host[5] = t.fit_rows;
    return kOk;
  });
}

int dpmi_sea_engine_set_tuning(dpmi_sea_engine* engine, const int* knobs_host) {
  return guarded([&] {
    if (!engine || !knobs_host) throw std::invalid_argument("null argument");
    dpmi::Tune& t = engine->impl->tune;
    const int* k = knobs_host;
    if (k[0] > 0) {
      if (k[0] % 64 || k[0] > 1024) throw std::invalid_argument("env threads per block: a multiple of 64 up to 1024");
      t.env_tpb = k[0];
    }
    if (k[1] > 0) {
      if (k[1] > 4) throw std::invalid_argument("forward embedding waves per workgroup: 1 to 4");
      t.fwd_waves = k[1];
    }
    if (k[2] > 0) {
      if (k[2] > 4) throw std::invalid_argument("adjoint embedding waves per workgroup: 1 to 4");
      t.bwd_waves = k[2];
    }
    if (k[3] > 0) {
      if (k[3] > 16 && k[3] != 32 && k[3] != 64 && k[3] != 128 && k[3] != 256)
        throw std::invalid_argument("force rows per workgroup: 1 to 16, 32, 64, 128 or 256");
      t.force_rows = k[3];
    }
    if (k[4] >= 0) {
      if (k[4] > 15) throw std::invalid_argument("matrix-core embedding mask: 0 to 15");
      t.embed_mfma = k[4];
    }
    if (k[5] > 0) {
      const int rows = k[5] & ~(dpmi::kFitSplit | dpmi::kFitWide);
      if ((rows != 32 && rows != 64) || ((k[5] & dpmi::kFitSplit) && (k[5] & dpmi::kFitWide)) ||
          ((k[5] & dpmi::kFitWide) && rows != 64))
        throw std::invalid_argument(
            "fused fitting rows per tile: 32 or 64, plus 256 for layer 0 in its own kernels, or 64 + 512 for 16 "
            "waves per tile");
      t.fit_rows = k[5];
    }
    return kOk;
  });
}

int dpmi_sea_engine_status(dpmi_sea_engine* engine) {
  return guarded([&] {
    if (!engine) throw std::invalid_argument("null engine");
    return flag_to_status(engine->impl->status_flags());
  });
}

int dpmi_sea_engine_fetch(dpmi_sea_engine* engine, int frame, int what, void* dst_host, long long* count_host) {
  return guarded([&] {
    const long long n = engine->impl->fetch(frame, what, dst_host);
    if (count_host) *count_host = n;
    return kOk;
  });
}

int dpmi_sea_engine_set_timing(dpmi_sea_engine* engine, int enable) {
  engine->impl->timing = enable != 0;
  engine->impl->reset_timing();
  return kOk;
}

int dpmi_sea_engine_timing(dpmi_sea_engine* engine, double* ms_host, int max_stages, char* names_host,
                           int names_len) {
  try {
    return engine->impl->timing_report(ms_host, max_stages, names_host, names_len);
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"
