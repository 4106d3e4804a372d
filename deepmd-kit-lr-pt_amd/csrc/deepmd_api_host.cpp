// deepmd::*_cpu drop-ins (include/deepmd_lib.h): the reference's host-array
// entry points, served by the same HIP kernels as the *_gpu templates.
//
// The library carries no CPU arithmetic.  Each function copies its host
// inputs to the current HIP device, runs the corresponding *_gpu path and
// copies the outputs back, so host and device callers get the same results
// (and the call fails loudly without a GPU).  Signatures, argument meaning,
// output layout and error behaviour follow source/lib/include:
//   normalize_coord_cpu, copy_coord_cpu          coord.h:9-37
//   build_nlist_cpu                              neighbor_list.h:166-176
//   format_nlist_cpu                             fmt_nlist.h:11-19
//   prod_env_mat_a_cpu                           prod_env_mat.h:10-49
//   tabulate_fusion_se_a{,_grad,_grad_grad}_cpu  tabulate.h:28-71
//   prod_force_a_cpu, prod_virial_a_cpu          prod_force.h:20-27, prod_virial.h:7-15
//   prod_{force,virial}_grad_a_cpu               prod_force_grad.h:7-13, prod_virial_grad.h:7-13
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/deepmd_lib.h"
#include "dpmi_common.h"
#include "kernels.h"

namespace deepmd {

namespace {

// A device copy of a host array, released at scope exit.
template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  explicit Dev(size_t count, const T* host = nullptr) : n(count) {
    if (hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
      throw deepmd_exception_oom("device staging allocation failed");
    if (host && n && hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(p);
      throw deepmd_exception("host to device copy failed");
    }
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  void down(T* host, size_t count) const {
    if (host && count && hipMemcpy(host, p, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
      throw deepmd_exception("device to host copy failed");
  }
};

// A device InputNlist (rows of width `width`) holding a host list.
struct DevNlist {
  InputNlist nl;
  int* rows = nullptr;
  int width = 0;
  DevNlist(const InputNlist& host, int width_) : width(std::max(width_, 1)) {
    if (hipMalloc((void**)&rows, sizeof(int) * std::max<size_t>((size_t)host.inum * width, 1)) != hipSuccess)
      throw deepmd_exception_oom("device staging allocation failed");
    InputNlist h = host;
    try {
      convert_nlist_gpu_device(nl, h, rows, width);
    } catch (...) {
      (void)hipFree(rows);
      throw;
    }
  }
  ~DevNlist() {
    free_nlist_gpu_device(nl);
    if (rows) (void)hipFree(rows);
  }
};

int max_row(const InputNlist& l) {
  int m = 0;
  for (int i = 0; i < l.inum; ++i) m = std::max(m, l.numneigh[i]);
  return m;
}

// rows of a table: int((upper - lower) / stride0 + (max - upper) / stride1)
template <typename T>
size_t table_rows(const T* info) {
  return (size_t)((info[1] - info[0]) / info[3] + (info[2] - info[1]) / info[4]);
}

}  // namespace

template <typename FPTYPE>
void normalize_coord_cpu(FPTYPE* coord, const int natom, const Region<FPTYPE>& region) {
  Dev<FPTYPE> c(3 * (size_t)natom, coord), b(9, region.boxt), r(9, region.rec_boxt);
  Region<FPTYPE> dr(b.p, r.p);
  normalize_coord_gpu<FPTYPE>(c.p, natom, dr);
  c.down(coord, 3 * (size_t)natom);
}

template <typename FPTYPE>
int copy_coord_cpu(FPTYPE* out_c, int* out_t, int* mapping, int* nall, const FPTYPE* in_c, const int* in_t,
                   const int& nloc, const int& mem_nall, const float& rcut, const Region<FPTYPE>& region) {
  int ci[23];
  compute_cell_info(ci, rcut, region);
  Dev<int> dci(23, ci);
  Dev<FPTYPE> ic(3 * (size_t)nloc, in_c), oc(3 * (size_t)mem_nall), b(9, region.boxt), r(9, region.rec_boxt);
  Dev<int> it(nloc, in_t), ot(mem_nall), om(mem_nall);
  Region<FPTYPE> dr(b.p, r.p);
  const int rc = copy_coord_gpu<FPTYPE>(oc.p, ot.p, om.p, nall, nullptr, ic.p, it.p, nloc, mem_nall, ci[21], ci[22],
                                        dci.p, dr);
  if (rc) return rc;
  oc.down(out_c, 3 * (size_t)*nall);
  ot.down(out_t, *nall);
  om.down(mapping, *nall);
  return 0;
}

template <typename FPTYPE>
int build_nlist_cpu(InputNlist& nlist, int* max_list_size, const FPTYPE* c_cpy, const int& nloc, const int& nall,
                    const int& mem_size, const float& rcut, const int& nframes, const int* type) {
  const size_t rows = (size_t)nframes * nloc;
  Dev<FPTYPE> c(3 * (size_t)nframes * nall, c_cpy);
  Dev<int> t(type ? (size_t)nframes * nall : 0, type);
  Dev<int> il(rows), nn(rows), jl(rows * std::max(mem_size, 1));
  try {
    dpmi::launch_brute_nlist<FPTYPE>(c.p, nloc, nall, mem_size, (FPTYPE)(rcut * rcut), nframes, type ? t.p : nullptr,
                                     il.p, nn.p, jl.p, nullptr);
  } catch (const dpmi::HipError& e) {
    throw deepmd_exception(e.what());
  }
  std::vector<int> h_nn(rows), h_jl(rows * std::max(mem_size, 1));
  nn.down(h_nn.data(), rows);
  il.down(nlist.ilist, rows);
  jl.down(h_jl.data(), h_jl.size());
  nlist.inum = (int)rows;
  int mx = 0;
  for (size_t r = 0; r < rows; ++r) {
    mx = std::max(mx, h_nn[r]);
    nlist.numneigh[r] = h_nn[r];
    const int k = std::min(h_nn[r], mem_size);
    std::copy(h_jl.begin() + r * mem_size, h_jl.begin() + r * mem_size + k, nlist.firstneigh[r]);
  }
  *max_list_size = mx;
  return mx > mem_size ? 1 : 0;
}

template <typename FPTYPE>
void format_nlist_cpu(int* nlist, const InputNlist& in_nlist, const FPTYPE* coord, const int* type, const int nloc,
                      const int nall, const float rcut, const std::vector<int> sec) {
  DevNlist g(in_nlist, max_row(in_nlist));
  Dev<FPTYPE> c(3 * (size_t)nall, coord);
  Dev<int> t(nall, type), out((size_t)nloc * sec.back());
  format_nbor_list_gpu<FPTYPE>(out.p, c.p, t.p, g.nl, nullptr, nullptr, g.width, nloc, nall, 1, rcut, sec);
  out.down(nlist, (size_t)nloc * sec.back());
}

template <typename FPTYPE>
void prod_env_mat_a_cpu(FPTYPE* em, FPTYPE* em_deriv, FPTYPE* rij, int* nlist, const FPTYPE* coord, const int* type,
                        const InputNlist& inlist, const int max_nbor_size, const FPTYPE* avg, const FPTYPE* std_,
                        const int nloc, const int nall, const int nframes, const float rcut, const float rcut_smth,
                        const std::vector<int> sec, const int* f_type) {
  const int ntypes = (int)sec.size() - 1;
  const size_t nnei = sec.back(), rows = (size_t)nframes * nloc;
  DevNlist g(inlist, std::max(max_nbor_size, max_row(inlist)));
  Dev<FPTYPE> c(3 * (size_t)nframes * nall, coord), a((size_t)ntypes * nnei * 4, avg),
      s((size_t)ntypes * nnei * 4, std_);
  Dev<int> t((size_t)nframes * nall, type), ft(f_type ? (size_t)nframes * nall : 0, f_type);
  Dev<FPTYPE> o_em(rows * nnei * 4), o_ed(rows * nnei * 12), o_r(rows * nnei * 3);
  Dev<int> o_nl(rows * nnei);
  prod_env_mat_a_gpu<FPTYPE>(o_em.p, o_ed.p, o_r.p, o_nl.p, c.p, t.p, g.nl, nullptr, nullptr, g.width, a.p, s.p, nloc,
                             nall, nframes, rcut, rcut_smth, sec, f_type ? ft.p : nullptr);
  o_em.down(em, rows * nnei * 4);
  o_ed.down(em_deriv, rows * nnei * 12);
  o_r.down(rij, rows * nnei * 3);
  o_nl.down(nlist, rows * nnei);
}

template <typename FPTYPE>
void tabulate_fusion_se_a_cpu(FPTYPE* out, const FPTYPE* table, const FPTYPE* table_info, const FPTYPE* em_x,
                              const FPTYPE* em, const FPTYPE* two_embed, const int nloc, const int nnei,
                              const int last_layer_size, const bool is_sorted, const int ndescrpt) {
  if (two_embed) throw deepmd_exception("two_embed (se_atten) is not part of the se_a path");
  const size_t M = last_layer_size;
  Dev<FPTYPE> tb(table_rows(table_info) * 6 * M, table), ex((size_t)nloc * nnei, em_x),
      e((size_t)nloc * nnei * ndescrpt, em), o((size_t)nloc * ndescrpt * M);
  tabulate_fusion_se_a_gpu<FPTYPE>(o.p, tb.p, table_info, ex.p, e.p, nullptr, nloc, nnei, last_layer_size, is_sorted,
                                   ndescrpt);
  o.down(out, (size_t)nloc * ndescrpt * M);
}

template <typename FPTYPE>
void tabulate_fusion_se_a_grad_cpu(FPTYPE* dy_dem_x, FPTYPE* dy_dem, FPTYPE* dy_dtwo, const FPTYPE* table,
                                   const FPTYPE* table_info, const FPTYPE* em_x, const FPTYPE* em,
                                   const FPTYPE* two_embed, const FPTYPE* dy, const int nloc, const int nnei,
                                   const int last_layer_size, const bool is_sorted, const int ndescrpt) {
  if (two_embed || dy_dtwo) throw deepmd_exception("two_embed (se_atten) is not part of the se_a path");
  const size_t M = last_layer_size;
  Dev<FPTYPE> tb(table_rows(table_info) * 6 * M, table), ex((size_t)nloc * nnei, em_x),
      e((size_t)nloc * nnei * ndescrpt, em), d((size_t)nloc * ndescrpt * M, dy), ox((size_t)nloc * nnei),
      oe((size_t)nloc * nnei * ndescrpt);
  tabulate_fusion_se_a_grad_gpu<FPTYPE>(ox.p, oe.p, nullptr, tb.p, table_info, ex.p, e.p, nullptr, d.p, nloc, nnei,
                                        last_layer_size, is_sorted, ndescrpt);
  ox.down(dy_dem_x, (size_t)nloc * nnei);
  oe.down(dy_dem, (size_t)nloc * nnei * ndescrpt);
}

template <typename FPTYPE>
void tabulate_fusion_se_a_grad_grad_cpu(FPTYPE* dz_dy, const FPTYPE* table, const FPTYPE* table_info,
                                        const FPTYPE* em_x, const FPTYPE* em, const FPTYPE* two_embed,
                                        const FPTYPE* dz_dy_dem_x, const FPTYPE* dz_dy_dem, const FPTYPE* dz_dy_dtwo,
                                        const int nloc, const int nnei, const int last_layer_size,
                                        const bool is_sorted, const int ndescrpt) {
  if (two_embed || dz_dy_dtwo) throw deepmd_exception("two_embed (se_atten) is not part of the se_a path");
  const size_t M = last_layer_size;
  Dev<FPTYPE> tb(table_rows(table_info) * 6 * M, table), ex((size_t)nloc * nnei, em_x),
      e((size_t)nloc * nnei * ndescrpt, em), zx((size_t)nloc * nnei, dz_dy_dem_x),
      ze((size_t)nloc * nnei * ndescrpt, dz_dy_dem), o((size_t)nloc * ndescrpt * M);
  tabulate_fusion_se_a_grad_grad_gpu<FPTYPE>(o.p, tb.p, table_info, ex.p, e.p, nullptr, zx.p, ze.p, nullptr, nloc,
                                             nnei, last_layer_size, is_sorted, ndescrpt);
  o.down(dz_dy, (size_t)nloc * ndescrpt * M);
}

template <typename FPTYPE>
void prod_force_a_cpu(FPTYPE* force, const FPTYPE* net_deriv, const FPTYPE* in_deriv, const int* nlist,
                      const int nloc, const int nall, const int nnei, const int nframes) {
  const size_t rows = (size_t)nframes * nloc;
  Dev<FPTYPE> nd(rows * nnei * 4, net_deriv), ed(rows * nnei * 12, in_deriv), f((size_t)nframes * nall * 3);
  Dev<int> nl(rows * nnei, nlist);
  prod_force_a_gpu<FPTYPE>(f.p, nd.p, ed.p, nl.p, nloc, nall, nnei, nframes);
  f.down(force, (size_t)nframes * nall * 3);
}

template <typename FPTYPE>
void prod_virial_a_cpu(FPTYPE* virial, FPTYPE* atom_virial, const FPTYPE* net_deriv, const FPTYPE* env_deriv,
                       const FPTYPE* rij, const int* nlist, const int nloc, const int nall, const int nnei) {
  Dev<FPTYPE> nd((size_t)nloc * nnei * 4, net_deriv), ed((size_t)nloc * nnei * 12, env_deriv),
      r((size_t)nloc * nnei * 3, rij), v(9), av((size_t)nall * 9);
  Dev<int> nl((size_t)nloc * nnei, nlist);
  prod_virial_a_gpu<FPTYPE>(v.p, av.p, nd.p, ed.p, r.p, nl.p, nloc, nall, nnei);
  v.down(virial, 9);
  av.down(atom_virial, (size_t)nall * 9);
}

template <typename FPTYPE>
void prod_force_grad_a_cpu(FPTYPE* grad_net, const FPTYPE* grad, const FPTYPE* env_deriv, const int* nlist,
                           const int nloc, const int nnei, const int nframes) {
  const size_t rows = (size_t)nframes * nloc;
  Dev<FPTYPE> g(rows * 3, grad), ed(rows * nnei * 12, env_deriv), o(rows * nnei * 4);
  Dev<int> nl(rows * nnei, nlist);
  prod_force_grad_a_gpu<FPTYPE>(o.p, g.p, ed.p, nl.p, nloc, nnei, nframes);
  o.down(grad_net, rows * nnei * 4);
}

template <typename FPTYPE>
void prod_virial_grad_a_cpu(FPTYPE* grad_net, const FPTYPE* grad, const FPTYPE* env_deriv, const FPTYPE* rij,
                            const int* nlist, const int nloc, const int nnei) {
  Dev<FPTYPE> g(9, grad), ed((size_t)nloc * nnei * 12, env_deriv), r((size_t)nloc * nnei * 3, rij),
      o((size_t)nloc * nnei * 4);
  Dev<int> nl((size_t)nloc * nnei, nlist);
  prod_virial_grad_a_gpu<FPTYPE>(o.p, g.p, ed.p, r.p, nl.p, nloc, nnei);
  o.down(grad_net, (size_t)nloc * nnei * 4);
}

#define DPMI_HOST_INST(T)                                                                                            \
  template void normalize_coord_cpu<T>(T*, const int, const Region<T>&);                                             \
  template int copy_coord_cpu<T>(T*, int*, int*, int*, const T*, const int*, const int&, const int&, const float&,   \
                                 const Region<T>&);                                                                 \
  template int build_nlist_cpu<T>(InputNlist&, int*, const T*, const int&, const int&, const int&, const float&,     \
                                  const int&, const int*);                                                           \
  template void format_nlist_cpu<T>(int*, const InputNlist&, const T*, const int*, const int, const int, const float, \
                                    const std::vector<int>);                                                         \
  template void prod_env_mat_a_cpu<T>(T*, T*, T*, int*, const T*, const int*, const InputNlist&, const int, const T*, \
                                      const T*, const int, const int, const int, const float, const float,           \
                                      const std::vector<int>, const int*);                                           \
  template void tabulate_fusion_se_a_cpu<T>(T*, const T*, const T*, const T*, const T*, const T*, const int,         \
                                            const int, const int, const bool, const int);                            \
  template void tabulate_fusion_se_a_grad_cpu<T>(T*, T*, T*, const T*, const T*, const T*, const T*, const T*,       \
                                                 const T*, const int, const int, const int, const bool, const int);  \
  template void tabulate_fusion_se_a_grad_grad_cpu<T>(T*, const T*, const T*, const T*, const T*, const T*,          \
                                                      const T*, const T*, const T*, const int, const int, const int, \
                                                      const bool, const int);                                        \
  template void prod_force_a_cpu<T>(T*, const T*, const T*, const int*, const int, const int, const int, const int); \
  template void prod_virial_a_cpu<T>(T*, T*, const T*, const T*, const T*, const int*, const int, const int,         \
                                     const int);                                                                     \
  template void prod_force_grad_a_cpu<T>(T*, const T*, const T*, const int*, const int, const int, const int);       \
  template void prod_virial_grad_a_cpu<T>(T*, const T*, const T*, const T*, const int*, const int, const int);
DPMI_HOST_INST(double)
DPMI_HOST_INST(float)

}  // namespace deepmd
