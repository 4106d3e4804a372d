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
#include <string>
#include <type_traits>
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
void normalize_coord_gpu<float>(float* coord, const int natom, const Region<float>& region) {
  // FP32 arithmetic on the Region's own box and reciprocal box (device rows)
  float b[9], r[9];
  if (hipMemcpy(b, region.boxt, sizeof(b), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(r, region.rec_boxt, sizeof(r), hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read the box of the region");
  guard([&] { dpmi::launch_normalize_f32(coord, natom, b, r, nullptr); });
}

namespace {

int copy_coord_double(double* out_c, int* out_t, int* mapping, int* nall, const double* in_c, const int* in_t,
                      int nloc, int mem_nall, const int* cell_info, const std::vector<double>& box) {
  // The grid (ncell, ngcell) is the caller's cell_info (coord.cc:69-108,
  // 23 ints on the device); the int_data workspace is not needed.
  int ci[23];
  if (hipMemcpy(ci, cell_info, sizeof(ci), hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read cell_info");
  dpmi::FrameGeom g;
  dpmi::geom_from_box(g, box.data(), 1.0f);
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

__global__ void k_to_double(const float* __restrict__ in, double* __restrict__ out, long long n) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (double)in[k];
}
__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, long long n) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (float)in[k];
}

}  // namespace

template <>
int copy_coord_gpu<double>(double* out_c, int* out_t, int* mapping, int* nall, int* /*int_data*/, const double* in_c,
                           const int* in_t, const int& nloc, const int& mem_nall, const int& /*loc_cellnum*/,
                           const int& /*total_cellnum*/, const int* cell_info, const Region<double>& region) {
  return copy_coord_double(out_c, out_t, mapping, nall, in_c, in_t, nloc, mem_nall, cell_info,
                           box_from_device(region.boxt));
}

template <>
int copy_coord_gpu<float>(float* out_c, int* out_t, int* mapping, int* nall, int* /*int_data*/, const float* in_c,
                          const int* in_t, const int& nloc, const int& mem_nall, const int& /*loc_cellnum*/,
                          const int& /*total_cellnum*/, const int* cell_info, const Region<float>& region) {
  // The ghost images are built in FP64 from the FP32 coordinates and box
  // (exact conversions) and rounded once to FP32.
  float bf[9];
  if (hipMemcpy(bf, region.boxt, sizeof(bf), hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read the box of the region");
  const std::vector<double> box(bf, bf + 9);
  double* in_d = static_cast<double*>(buf(2).get(sizeof(double) * 3 * (size_t)std::max(nloc, 1)));
  double* out_d = static_cast<double*>(buf(3).get(sizeof(double) * 3 * (size_t)std::max(mem_nall, 1)));
  guard([&] {
    if (nloc > 0)
      hipLaunchKernelGGL(k_to_double, dim3(dpmi::ceil_div(3LL * nloc, 256)), dim3(256), 0, nullptr, in_c, in_d,
                         3LL * nloc);
    DPMI_LAUNCH_CHECK();
  });
  const int rc = copy_coord_double(out_d, out_t, mapping, nall, in_d, in_t, nloc, mem_nall, cell_info, box);
  const long long n = 3LL * std::min(*nall, mem_nall);
  guard([&] {
    if (n > 0) hipLaunchKernelGGL(k_to_float, dim3(dpmi::ceil_div(n, 256)), dim3(256), 0, nullptr, out_d, out_c, n);
    DPMI_LAUNCH_CHECK();
  });
  return rc;
}

template <typename FPTYPE>
int build_nlist_gpu(InputNlist& nlist, int* max_list_size, int* /*nlist_data*/, const FPTYPE* c_cpy, const int& nloc,
                    const int& nall, const int& mem_size, const float& rcut, const int& nframes, const int* type) {
  // neighbor_list.cu:198-250 contract: mem_size must cover nall
  if (mem_size < nall) return 1;
  guard([&] {
    dpmi::launch_brute_nlist<FPTYPE>(c_cpy, nloc, nall, mem_size, (FPTYPE)(rcut * rcut), nframes, type, nlist.ilist,
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
template int build_nlist_gpu<double>(InputNlist&, int*, int*, const double*, const int&, const int&, const int&,
                                     const float&, const int&, const int*);
template int build_nlist_gpu<float>(InputNlist&, int*, int*, const float*, const int&, const int&, const int&,
                                    const float&, const int&, const int*);

namespace {

// Formats the device InputNlist rows into nlist [nframes * nloc, sec.back()]
// (-1 padding); returns the device status bits.
template <typename FPTYPE>
int format_rows(int* nlist, const FPTYPE* coord, const int* type, const InputNlist& gpu_inlist, int max_nbor_size,
                int nloc, int nall, int nframes, float rcut, const std::vector<int>& sec) {
  const int ntypes = (int)sec.size() - 1;
  if (ntypes <= 0 || ntypes > 127) throw deepmd_exception("the number of types must be in [1, 127]");
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
    dpmi::launch_row_format<FPTYPE>(gpu_inlist.inum, gpu_inlist.ilist, gpu_inlist.numneigh, nullptr, 0, coord, type,
                                    nloc, nall, na, flag, nullptr, gpu_inlist.firstneigh);
  });
  int f = 0;
  if (hipMemcpy(&f, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) throw deepmd_exception("flag read");
  return f;
}

void throw_format_flags(int f) {
  if (f & dpmi::kFlagNlistCapacity) throw deepmd_exception_nlist_capacity("neighbour row exceeds max_nbor_size");
  if (f) throw deepmd_exception("Illegal nbor list sorting");
}

}  // namespace

template <typename FPTYPE>
void format_nbor_list_gpu(int* nlist, const FPTYPE* coord, const int* type, const InputNlist& gpu_inlist,
                          int* /*array_int*/, unsigned long long* /*array_longlong*/, const int max_nbor_size,
                          const int nloc, const int nall, const int nframes, const float rcut,
                          const std::vector<int> sec) {
  throw_format_flags(format_rows(nlist, coord, type, gpu_inlist, max_nbor_size, nloc, nall, nframes, rcut, sec));
}
template void format_nbor_list_gpu<double>(int*, const double*, const int*, const InputNlist&, int*,
                                           unsigned long long*, const int, const int, const int, const int,
                                           const float, const std::vector<int>);
template void format_nbor_list_gpu<float>(int*, const float*, const int*, const InputNlist&, int*,
                                          unsigned long long*, const int, const int, const int, const int,
                                          const float, const std::vector<int>);

template <typename FPTYPE>
void prod_env_mat_a_gpu(FPTYPE* em, FPTYPE* em_deriv, FPTYPE* rij, int* nlist, const FPTYPE* coord, const int* type,
                        const InputNlist& gpu_inlist, int* /*array_int*/, unsigned long long* /*array_longlong*/,
                        const int max_nbor_size, const FPTYPE* avg, const FPTYPE* std_, const int nloc,
                        const int nall, const int nframes, const float rcut, const float rcut_smth,
                        const std::vector<int> sec, const int* f_type) {
  if (!f_type) f_type = type;
  const int ntypes = (int)sec.size() - 1;
  const int nnei = sec.back();
  const int f = format_rows(nlist, coord, f_type, gpu_inlist, max_nbor_size, nloc, nall, nframes, rcut, sec);
  throw_format_flags(f);
  dpmi::EnvArgs ea;
  ea.type = type;
  ea.nlist = nlist;
  ea.strided_nloc = nloc;
  ea.nrows = (long long)nframes * nloc;
  ea.nnei = nnei;
  ea.nall_stride = nall;
  ea.ntypes = ntypes;
  ea.rmin = (double)rcut_smth;
  ea.rmax = (double)rcut;
  if constexpr (std::is_same_v<FPTYPE, double>) {
    ea.coord = coord;
    ea.avg = avg;
    ea.std_ = std_;
    ea.em = em;
    ea.em_deriv = em_deriv;
    ea.rij = rij;
    guard([&] { dpmi::launch_env_mat_a(ea, nullptr); });
  } else {
    // FP32 rows: the geometry is evaluated in FP64 from the FP32 coordinates
    // and statistics (exact conversions) and each stored value rounded once
    const size_t ns = (size_t)ntypes * nnei * 4;
    double* st = static_cast<double*>(buf(0).get(sizeof(double) * 2 * std::max<size_t>(ns, 1)));
    guard([&] {
      hipLaunchKernelGGL(k_to_double, dim3(dpmi::ceil_div((long long)ns, 256)), dim3(256), 0, nullptr, avg, st,
                         (long long)ns);
      hipLaunchKernelGGL(k_to_double, dim3(dpmi::ceil_div((long long)ns, 256)), dim3(256), 0, nullptr, std_, st + ns,
                         (long long)ns);
      DPMI_LAUNCH_CHECK();
      ea.coord_f = coord;
      ea.avg = st;
      ea.std_ = st + ns;
      dpmi::launch_env_mat_a_f32(ea, em, em_deriv, rij, nullptr);
    });
  }
}
template void prod_env_mat_a_gpu<double>(double*, double*, double*, int*, const double*, const int*,
                                         const InputNlist&, int*, unsigned long long*, const int, const double*,
                                         const double*, const int, const int, const int, const float, const float,
                                         const std::vector<int>, const int*);
template void prod_env_mat_a_gpu<float>(float*, float*, float*, int*, const float*, const int*, const InputNlist&,
                                        int*, unsigned long long*, const int, const float*, const float*, const int,
                                        const int, const int, const float, const float, const std::vector<int>,
                                        const int*);

// coord.cc:69-108 (host Region, as the reference computes it on the host)
template <typename FPTYPE>
void compute_cell_info(int* cell_info, const float& rcut, const Region<FPTYPE>& region) {
  double b[9];
  for (int k = 0; k < 9; ++k) b[k] = region.boxt[k];
  dpmi::FrameGeom g;
  dpmi::geom_from_box(g, b, rcut);
  // geom_from_box keeps the face distances implicit: recompute the cell sizes
  // from the same grid (cell_size = to_face / ncell) for cell_iter
  double tf[3];
  dpmi::geom_face_distances(b, tf);
  for (int d = 0; d < 3; ++d) {
    const double cs = tf[d] / g.nc[d];
    cell_info[d] = 0;
    cell_info[3 + d] = g.nc[d];
    cell_info[12 + d] = g.ng[d];
    cell_info[6 + d] = -g.ng[d];
    cell_info[9 + d] = g.nc[d] + g.ng[d];
    cell_info[15 + d] = g.ng[d];
    int it = (int)(rcut / cs);
    if (it * cs < rcut) it += 1;
    cell_info[18 + d] = it;
  }
  cell_info[21] = g.nc[0] * g.nc[1] * g.nc[2];
  if (cell_info[21] <= 0)
    throw deepmd_exception("loc_cellnum should be positive but is " + std::to_string(cell_info[21]) +
                           ". You may give a PBC box with zero volume.");
  cell_info[22] = g.nx[0] * g.nx[1] * g.nx[2];
}
template void compute_cell_info<double>(int*, const float&, const Region<double>&);
template void compute_cell_info<float>(int*, const float&, const Region<float>&);

// neighbor_list.cc:993-1008
int validate_nlist_gpu_capacity(const InputNlist& nlist, const int row_capacity) {
  int mx = 0;
  for (int ii = 0; ii < nlist.inum; ++ii) {
    const int n = nlist.numneigh[ii];
    if (n > row_capacity)
      throw deepmd_exception_nlist_capacity("Neighbor list row " + std::to_string(ii) + " contains " +
                                            std::to_string(n) +
                                            " neighbors, exceeding the allocated GPU row capacity " +
                                            std::to_string(row_capacity));
    mx = std::max(mx, n);
  }
  return mx;
}

namespace {
template <typename T>
T* dev_alloc(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw deepmd_exception_oom("hipMalloc failed");
  return static_cast<T*>(p);
}
void dev_free(void* p) {
  if (p) (void)hipFree(p);
}
void h2d(void* d, const void* h, size_t bytes) {
  if (bytes && hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess)
    throw deepmd_exception("host to device copy failed");
}
}  // namespace

// neighbor_list.cc:1010-1039: rows of the host list at gpu_memory + ii *
// max_nbor_size (the caller's device buffer), device ilist / numneigh /
// firstneigh allocated here and released by free_nlist_gpu_device.
void convert_nlist_gpu_device(InputNlist& gpu_nlist, InputNlist& cpu_nlist, int*& gpu_memory,
                              const int& max_nbor_size) {
  validate_nlist_gpu_capacity(cpu_nlist, max_nbor_size);
  const int inum = cpu_nlist.inum;
  gpu_nlist.inum = inum;
  gpu_nlist.ilist = dev_alloc<int>(inum);
  gpu_nlist.numneigh = dev_alloc<int>(inum);
  gpu_nlist.firstneigh = dev_alloc<int*>(inum);
  h2d(gpu_nlist.ilist, cpu_nlist.ilist, sizeof(int) * inum);
  h2d(gpu_nlist.numneigh, cpu_nlist.numneigh, sizeof(int) * inum);
  std::vector<int*> rows(inum);
  std::vector<int> packed((size_t)inum * max_nbor_size, 0);
  for (int ii = 0; ii < inum; ++ii) {
    std::copy(cpu_nlist.firstneigh[ii], cpu_nlist.firstneigh[ii] + cpu_nlist.numneigh[ii],
              packed.begin() + (size_t)ii * max_nbor_size);
    rows[ii] = gpu_memory + (size_t)ii * max_nbor_size;
  }
  h2d(gpu_memory, packed.data(), sizeof(int) * packed.size());
  h2d(gpu_nlist.firstneigh, rows.data(), sizeof(int*) * inum);
}

void free_nlist_gpu_device(InputNlist& gpu_nlist) {
  dev_free(gpu_nlist.ilist);
  dev_free(gpu_nlist.numneigh);
  dev_free(gpu_nlist.firstneigh);
  gpu_nlist.ilist = nullptr;
  gpu_nlist.numneigh = nullptr;
  gpu_nlist.firstneigh = nullptr;
}

namespace {
__global__ void k_map_nlist(int* __restrict__ nlist, const int* __restrict__ nlist_map, long long n) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int j = nlist[k];
  if (j != -1) nlist[k] = nlist_map[j];
}
}  // namespace

// neighbor_list.cu:252-264: nlist entries mapped through nlist_map
void use_nlist_map(int* nlist, const int* nlist_map, const int nloc, const int nnei) {
  const long long n = (long long)nloc * nnei;
  guard([&] {
    if (n > 0) hipLaunchKernelGGL(k_map_nlist, dim3(dpmi::ceil_div(n, 256)), dim3(256), 0, nullptr, nlist, nlist_map, n);
    DPMI_LAUNCH_CHECK();
  });
}

// prod_env_mat.cc:319-385: the host InputNlist whose pointers are packed in
// mesh[4..15] (device array of `size` ints, mesh[0] = ago) is uploaded when
// ago == 0 or it has more rows than the device copy: rows of width
// max_nbor_size (256, 512, ..., 4096) in nbor_list_dev, device ilist /
// numneigh / firstneigh in gpu_inlist (grown as needed).
void env_mat_nbor_update(InputNlist& inlist, InputNlist& gpu_inlist, int& max_nbor_size, int*& nbor_list_dev,
                         const int* mesh, const int size) {
  std::vector<int> mh(size);
  if (size > 0 && hipMemcpy(mh.data(), mesh, sizeof(int) * size, hipMemcpyDeviceToHost) != hipSuccess)
    throw deepmd_exception("cannot read mesh");
  if (size < 16) throw deepmd_exception("mesh must hold 16 ints for an external neighbour list");
  std::memcpy(&inlist.ilist, 4 + mh.data(), sizeof(int*));
  std::memcpy(&inlist.numneigh, 8 + mh.data(), sizeof(int*));
  std::memcpy(&inlist.firstneigh, 12 + mh.data(), sizeof(int**));
  const int ago = mh[0];
  if (ago != 0 && gpu_inlist.inum >= inlist.inum) return;
  const int inum = inlist.inum;
  int width = validate_nlist_gpu_capacity(inlist, 4096);
  width = width <= 256 ? 256 : width <= 512 ? 512 : width <= 1024 ? 1024 : width <= 2048 ? 2048 : 4096;
  if (gpu_inlist.inum < inum) {
    free_nlist_gpu_device(gpu_inlist);
    gpu_inlist.ilist = dev_alloc<int>(inum);
    gpu_inlist.numneigh = dev_alloc<int>(inum);
    gpu_inlist.firstneigh = dev_alloc<int*>(inum);
  }
  h2d(gpu_inlist.ilist, inlist.ilist, sizeof(int) * inum);
  h2d(gpu_inlist.numneigh, inlist.numneigh, sizeof(int) * inum);
  if (nbor_list_dev == nullptr || width > max_nbor_size || inum > gpu_inlist.inum) {
    dev_free(nbor_list_dev);
    nbor_list_dev = dev_alloc<int>((size_t)inum * width);
  }
  gpu_inlist.inum = inum;
  max_nbor_size = width;
  std::vector<int> packed((size_t)inum * max_nbor_size, 0);
  std::vector<int*> rows(inum);
  for (int ii = 0; ii < inum; ++ii) {
    rows[ii] = nbor_list_dev + (size_t)ii * max_nbor_size;
    std::copy(inlist.firstneigh[ii], inlist.firstneigh[ii] + inlist.numneigh[ii],
              packed.begin() + (size_t)ii * max_nbor_size);
  }
  h2d(nbor_list_dev, packed.data(), sizeof(int) * packed.size());
  h2d(gpu_inlist.firstneigh, rows.data(), sizeof(int*) * inum);
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
