// deepmd:: drop-in templates (include/deepmd_lib.h) over the dpmi kernels.
//
// Each function keeps the reference's argument meaning and completion
// semantics: it runs on the default stream and returns after the device has
// finished (the reference synchronises after every launch), throws
// deepmd_exception on HIP failures and reports capacity failures the way the
// reference does.
#include "../../include/deepmd_lib.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/dpmi.h"
#include "dpmi_common.h"
#include "geometry.h"
#include "kernels.h"

namespace deepmd {

namespace {

void sync_or_throw() {
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) throw deepmd_exception(std::string("CUDA Assert: ") + hipGetErrorString(e));
}

void ok_or_throw(int rc) {
  if (rc == DPMI_OK) return;
  if (rc == DPMI_ERR_NLIST_CAPACITY) throw deepmd_exception_nlist_capacity(dpmi_last_error());
  if (rc == DPMI_ERR_MEMORY) throw deepmd_exception_oom(dpmi_last_error());
  throw deepmd_exception(dpmi_last_error());
}

template <class F>
void guard(F&& f) {
  try {
    f();
  } catch (const dpmi::HipError& e) {
    throw deepmd_exception(e.what());
  }
  sync_or_throw();
}

std::vector<double> box_from_device(const double* boxt_dev) {
  std::vector<double> b(9);
  if (hipMemcpy(b.data(), boxt_dev, sizeof(double) * 9, hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read the box of the region");
  return b;
}

// scratch for the virial partials and the pointer-row formatting
struct Buf {
  void* p = nullptr;
  size_t n = 0;
  void* get(size_t bytes) {
    if (bytes > n) {
      if (p) (void)hipFree(p);
      p = nullptr;
      if (hipMalloc(&p, bytes) != hipSuccess) throw deepmd_exception_oom("scratch allocation failed");
      n = bytes;
    }
    return p;
  }
};
Buf& buf(int k) {
  static Buf* b = new Buf[4];
  return b[k];
}

}  // namespace

// region.cc:12-31, 104-108
template <typename FPTYPE>
Region<FPTYPE>::Region() {
  boxt = new FPTYPE[9];
  rec_boxt = new FPTYPE[9];
  self_allocated = true;
}
template <typename FPTYPE>
Region<FPTYPE>::Region(FPTYPE* extern_boxt, FPTYPE* extern_rec_boxt) {
  boxt = extern_boxt;
  rec_boxt = extern_rec_boxt;
  self_allocated = false;
}
template <typename FPTYPE>
Region<FPTYPE>::~Region() {
  if (self_allocated) {
    delete[] boxt;
    delete[] rec_boxt;
  }
}
template struct Region<double>;
template struct Region<float>;

template <typename FPTYPE>
void init_region_cpu(Region<FPTYPE>& region, const FPTYPE* boxt) {
  double b[9];
  for (int k = 0; k < 9; ++k) b[k] = boxt[k];
  dpmi::FrameGeom g;
  dpmi::geom_from_box(g, b, 1.0f);
  for (int k = 0; k < 9; ++k) {
    region.boxt[k] = boxt[k];
    region.rec_boxt[k] = (FPTYPE)g.rec[k];
  }
}
template void init_region_cpu<double>(Region<double>&, const double*);
template void init_region_cpu<float>(Region<float>&, const float*);

template <>
void normalize_coord_gpu<double>(double* coord, const int natom, const Region<double>& region) {
  const std::vector<double> b = box_from_device(region.boxt);
  ok_or_throw(dpmi_normalize_coord_f64(coord, natom, b.data(), nullptr));
  sync_or_throw();
}

template <>
int copy_coord_gpu<double>(double* out_c, int* out_t, int* mapping, int* nall, int* /*int_data*/, const double* in_c,
                           const int* in_t, const int& nloc, const int& mem_nall, const int& /*loc_cellnum*/,
                           const int& /*total_cellnum*/, const int* cell_info, const Region<double>& region) {
  // The grid (ncell, ngcell) is the caller's cell_info (coord.cc:69-108,
  // 23 ints on the device); the int_data workspace is not needed.
  const std::vector<double> b = box_from_device(region.boxt);
  int ci[23];
  if (hipMemcpy(ci, cell_info, sizeof(ci), hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read cell_info");
  dpmi::FrameGeom g;
  dpmi::geom_from_box(g, b.data(), 1.0f);
  for (int d = 0; d < 3; ++d) {
    g.nc[d] = ci[3 + d];
    g.ng[d] = ci[12 + d];
    g.nx[d] = g.nc[d] + 2 * g.ng[d];
  }
  int n = 0;
  int rcode;
  try {
    rcode = dpmi::copy_coord_run(g, out_c, out_t, mapping, &n, in_c, in_t, nloc, mem_nall, nullptr);
  } catch (const std::exception& e) {
    throw deepmd_exception(e.what());
  }
  *nall = n;
  if (rcode == DPMI_ERR_MEMORY) return 1;
  ok_or_throw(rcode);
  return 0;
}

template <>
int build_nlist_gpu<double>(InputNlist& nlist, int* max_list_size, int* /*nlist_data*/, const double* c_cpy,
                            const int& nloc, const int& nall, const int& mem_size, const float& rcut,
                            const int& nframes, const int* type) {
  // neighbor_list.cu:198-250 contract: mem_size must cover nall
  if (mem_size < nall) return 1;
  guard([&] {
    dpmi::launch_brute_nlist(c_cpy, nloc, nall, mem_size, (double)(rcut * rcut), nframes, type, nlist.ilist,
                             nlist.numneigh, nullptr, nullptr, nlist.firstneigh);
  });
  nlist.inum = nframes * nloc;
  std::vector<int> nn(nlist.inum);
  if (nlist.inum && hipMemcpy(nn.data(), nlist.numneigh, sizeof(int) * nn.size(), hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read numneigh");
  int mx = 0;
  for (int v : nn) mx = std::max(mx, v);
  *max_list_size = mx;
  return 0;
}

template <>
void prod_env_mat_a_gpu<double>(double* em, double* em_deriv, double* rij, int* nlist, const double* coord,
                                const int* type, const InputNlist& gpu_inlist, int* /*array_int*/,
                                unsigned long long* /*array_longlong*/, const int max_nbor_size, const double* avg,
                                const double* std_, const int nloc, const int nall, const int nframes,
                                const float rcut, const float rcut_smth, const std::vector<int> sec,
                                const int* f_type) {
  if (!f_type) f_type = type;
  const int ntypes = (int)sec.size() - 1;
  const int nnei = sec.back();
  int* d_sec = static_cast<int*>(buf(1).get(sizeof(int) * (sec.size() + 1)));
  int* flag = d_sec + sec.size();
  guard([&] {
    DPMI_CHECK(hipMemcpy(d_sec, sec.data(), sizeof(int) * sec.size(), hipMemcpyHostToDevice));
    DPMI_CHECK(hipMemset(flag, 0, sizeof(int)));
    DPMI_CHECK(hipMemset(nlist, 0xff, sizeof(int) * (size_t)nframes * nloc * nnei));
    dpmi::NlistArgs na;
    na.cap = dpmi::nlist_row_capacity(std::max(max_nbor_size, 64));
    na.sec = d_sec;
    na.ntypes = ntypes;
    na.nnei = nnei;
    na.rc2f = rcut * rcut;
    na.nlist = nlist;
    dpmi::launch_row_format(gpu_inlist.inum, gpu_inlist.ilist, gpu_inlist.numneigh, nullptr, 0, coord, f_type, nloc,
                            nall, na, flag, nullptr, gpu_inlist.firstneigh);
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
    dpmi::launch_env_mat_a(ea, nullptr);
  });
  int f = 0;
  if (hipMemcpy(&f, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) throw deepmd_exception("flag read");
  if (f & dpmi::kFlagNlistCapacity) throw deepmd_exception_nlist_capacity("neighbour row exceeds max_nbor_size");
  if (f) throw deepmd_exception("Illegal nbor list sorting");
}

#define DPMI_TAB_TEMPLATES(T, SUF)                                                                                    \
  template <>                                                                                                         \
  void tabulate_fusion_se_a_gpu<T>(T * out, const T* table, const T* table_info, const T* em_x, const T* em,          \
                                   const T* two_embed, const int nloc, const int nnei, const int M,                  \
                                   const bool is_sorted, const int ndescrpt) {                                        \
    if (two_embed) throw deepmd_exception("two_embed (se_atten) is not part of the se_a path");                        \
    ok_or_throw(dpmi_tabulate_fusion_se_a_##SUF(out, table, table_info, em_x, em, nloc, nnei, M, is_sorted, ndescrpt, \
                                               nullptr));                                                             \
    sync_or_throw();                                                                                                  \
  }                                                                                                                   \
  template <>                                                                                                         \
  void tabulate_fusion_se_a_grad_gpu<T>(T * dy_dem_x, T * dy_dem, T * dy_dtwo, const T* table, const T* table_info,   \
                                        const T* em_x, const T* em, const T* two_embed, const T* dy, const int nloc,  \
                                        const int nnei, const int M, const bool is_sorted, const int ndescrpt) {      \
    if (two_embed || dy_dtwo) throw deepmd_exception("two_embed (se_atten) is not part of the se_a path");            \
    if (nloc > 0 && nnei > 0) {                                                                                       \
      (void)hipMemset(dy_dem_x, 0, sizeof(T) * (size_t)nloc * nnei);                                                  \
      (void)hipMemset(dy_dem, 0, sizeof(T) * (size_t)nloc * nnei * ndescrpt);                                         \
    }                                                                                                                 \
    ok_or_throw(dpmi_tabulate_fusion_se_a_grad_##SUF(dy_dem_x, dy_dem, table, table_info, em_x, em, dy, nloc, nnei,  \
                                                    M, is_sorted, ndescrpt, nullptr));                                \
    sync_or_throw();                                                                                                  \
  }                                                                                                                   \
  template <>                                                                                                         \
  void tabulate_fusion_se_a_grad_grad_gpu<T>(T * dz_dy, const T* table, const T* table_info, const T* em_x,          \
                                             const T* em, const T* two_embed, const T* dz_dy_dem_x,                   \
                                             const T* dz_dy_dem, const T* dz_dy_dtwo, const int nloc,                 \
                                             const int nnei, const int M, const bool is_sorted,                       \
                                             const int ndescrpt) {                                                    \
    if (two_embed || dz_dy_dtwo) throw deepmd_exception("two_embed (se_atten) is not part of the se_a path");         \
    ok_or_throw(dpmi_tabulate_fusion_se_a_grad_grad_##SUF(dz_dy, table, table_info, em_x, em, dz_dy_dem_x, dz_dy_dem, \
                                                         nloc, nnei, M, is_sorted, ndescrpt, nullptr));               \
    sync_or_throw();                                                                                                  \
  }

DPMI_TAB_TEMPLATES(double, f64)
DPMI_TAB_TEMPLATES(float, f32)

#define DPMI_FORCE_TEMPLATES(T, SUF)                                                                            \
  template <>                                                                                                   \
  void prod_force_a_gpu<T>(T * force, const T* net_deriv, const T* in_deriv, const int* nlist, const int nloc,  \
                           const int nall, const int nnei, const int nframes) {                                 \
    ok_or_throw(dpmi_prod_force_a_##SUF(force, net_deriv, in_deriv, nlist, nloc, nall, nnei, nframes, nullptr)); \
    sync_or_throw();                                                                                            \
  }                                                                                                             \
  template <>                                                                                                   \
  void prod_virial_a_gpu<T>(T * virial, T * atom_virial, const T* net_deriv, const T* env_deriv, const T* rij,  \
                            const int* nlist, const int nloc, const int nall, const int nnei) {                 \
    ok_or_throw(                                                                                                \
        dpmi_prod_virial_a_##SUF(virial, atom_virial, net_deriv, env_deriv, rij, nlist, nloc, nall, nnei, nullptr)); \
    sync_or_throw();                                                                                            \
  }                                                                                                             \
  template <>                                                                                                   \
  void prod_force_grad_a_gpu<T>(T * grad_net, const T* grad, const T* env_deriv, const int* nlist,              \
                                const int nloc, const int nnei, const int nframes) {                            \
    ok_or_throw(dpmi_prod_force_grad_a_##SUF(grad_net, grad, env_deriv, nlist, nloc, nnei, nframes, nullptr));  \
    sync_or_throw();                                                                                            \
  }                                                                                                             \
  template <>                                                                                                   \
  void prod_virial_grad_a_gpu<T>(T * grad_net, const T* grad, const T* env_deriv, const T* rij,                 \
                                 const int* nlist, const int nloc, const int nnei) {                            \
    ok_or_throw(dpmi_prod_virial_grad_a_##SUF(grad_net, grad, env_deriv, rij, nlist, nloc, nnei, nullptr));     \
    sync_or_throw();                                                                                            \
  }

DPMI_FORCE_TEMPLATES(double, f64)
DPMI_FORCE_TEMPLATES(float, f32)

}  // namespace deepmd
