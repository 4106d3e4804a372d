// C++ caller of the drop-in library, used by tests/test_gpu_cpp_harness.py.
//
//   dpmi_harness lib <dir>      every deepmd:: template of include/deepmd_lib.h
//                               called as the reference's callers call them
//   dpmi_harness deeppot <dir>  dpmi::DeepPot (include/dpmi_deeppot.h)
//
// Inputs and outputs are raw little-endian arrays in <dir>, one file per
// array (name.f64 / name.f32 / name.i32); the test writes the inputs and
// checks the outputs against the oracle and the reference's goldens.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "deepmd_lib.h"
#include "dpmi_deeppot.h"

namespace {

std::string g_dir;

template <typename T>
std::vector<T> rd(const std::string& name) {
  std::ifstream f(g_dir + "/" + name, std::ios::binary);
  if (!f) throw std::runtime_error("missing input " + name);
  std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<T> v(b.size() / sizeof(T));
  std::memcpy(v.data(), b.data(), v.size() * sizeof(T));
  return v;
}
template <typename T>
void wr(const std::string& name, const T* p, size_t n) {
  std::ofstream f(g_dir + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), n * sizeof(T));
}
template <typename T>
void wr(const std::string& name, const std::vector<T>& v) {
  wr(name, v.data(), v.size());
}

void hk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
template <typename T>
T* dnew(size_t n, const T* host = nullptr) {
  T* p = nullptr;
  hk(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
  if (host && n) hk(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice), "h2d");
  return p;
}
template <typename T>
std::vector<T> dget(const T* p, size_t n) {
  std::vector<T> v(n);
  if (n) hk(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "d2h");
  return v;
}

// host InputNlist over rows kept in `store`
struct HostNlist {
  std::vector<int> ilist, numneigh;
  std::vector<std::vector<int>> rows;
  std::vector<int*> first;
  deepmd::InputNlist nl;
  void make() {
    first.resize(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) first[i] = rows[i].data();
    nl = deepmd::InputNlist((int)ilist.size(), ilist.data(), numneigh.data(), first.data());
  }
};

template <typename T>
struct System {
  int nloc, ntypes, nnei;
  float rcut, rcut_smth;
  std::vector<int> sel, sec, atype;
  std::vector<T> coord, box, avg, std_;
};

template <typename T>
System<T> load_system() {
  System<T> s;
  const std::vector<int> p = rd<int>("params.i32");
  s.nloc = p[0];
  s.ntypes = p[1];
  s.sel.assign(p.begin() + 2, p.begin() + 2 + s.ntypes);
  s.sec.assign(1, 0);
  for (int v : s.sel) s.sec.push_back(s.sec.back() + v);
  s.nnei = s.sec.back();
  const std::vector<double> pf = rd<double>("params.f64");
  s.rcut = (float)pf[0];
  s.rcut_smth = (float)pf[1];
  s.atype = rd<int>("atype.i32");
  for (auto* dst : {&s.coord, &s.box, &s.avg, &s.std_}) dst->clear();
  for (double v : rd<double>("coord.f64")) s.coord.push_back((T)v);
  for (double v : rd<double>("box.f64")) s.box.push_back((T)v);
  for (double v : rd<double>("avg.f64")) s.avg.push_back((T)v);
  for (double v : rd<double>("std.f64")) s.std_.push_back((T)v);
  return s;
}

// The geometry and environment chain of the reference's ROCm op
// (prod_env_mat_multi_device.cc:_norm_copy_coord_gpu / _build_nlist_gpu /
// prod_env_mat_a_gpu) on device arrays; outputs prefixed by `pre`.
template <typename T>
void geometry_chain(const std::string& pre, System<T>& s, std::vector<T>& ext_c, std::vector<int>& ext_t,
                    std::vector<int>& mapping, HostNlist& hn) {
  const char* sfx = sizeof(T) == 8 ? ".f64" : ".f32";
  deepmd::Region<T> hreg;
  deepmd::init_region_cpu(hreg, s.box.data());
  T* d_boxt = dnew<T>(9, hreg.boxt);
  T* d_rec = dnew<T>(9, hreg.rec_boxt);
  deepmd::Region<T> dreg(d_boxt, d_rec);
  T* d_coord = dnew<T>(3 * s.nloc, s.coord.data());
  deepmd::normalize_coord_gpu<T>(d_coord, s.nloc, dreg);
  const std::vector<T> norm = dget(d_coord, 3 * s.nloc);
  wr(pre + "norm" + sfx, norm);
  int ci[23];
  deepmd::compute_cell_info(ci, s.rcut, hreg);
  wr(pre + "cellinfo.i32", ci, 23);
  int* d_ci = dnew<int>(23, ci);
  int* d_type = dnew<int>(s.nloc, s.atype.data());
  const int mem_nall = 64 * s.nloc;
  T* d_ext = dnew<T>(3 * (size_t)mem_nall);
  int *d_ext_t = dnew<int>(mem_nall), *d_map = dnew<int>(mem_nall);
  int nall = 0;
  const int rc = deepmd::copy_coord_gpu<T>(d_ext, d_ext_t, d_map, &nall, nullptr, d_coord, d_type, s.nloc, mem_nall,
                                           ci[21], ci[22], d_ci, dreg);
  if (rc) throw std::runtime_error("copy_coord_gpu: mem_nall too small");
  ext_c = dget(d_ext, 3 * (size_t)nall);
  ext_t = dget(d_ext_t, nall);
  mapping = dget(d_map, nall);
  wr(pre + "ext_coord" + sfx, ext_c);
  wr(pre + "ext_type.i32", ext_t);
  wr(pre + "mapping.i32", mapping);
  // build_nlist_gpu with a device InputNlist (rows of width mem_size = nall)
  const int mem = nall;
  int *d_il = dnew<int>(s.nloc), *d_nn = dnew<int>(s.nloc), *d_rows = dnew<int>((size_t)s.nloc * mem);
  std::vector<int*> rp(s.nloc);
  for (int i = 0; i < s.nloc; ++i) rp[i] = d_rows + (size_t)i * mem;
  int** d_fn = dnew<int*>(s.nloc, rp.data());
  deepmd::InputNlist gl(0, d_il, d_nn, d_fn);
  int max_size = 0;
  if (deepmd::build_nlist_gpu<T>(gl, &max_size, nullptr, d_ext, s.nloc, nall, mem, s.rcut, 1, d_ext_t))
    throw std::runtime_error("build_nlist_gpu capacity");
  hn.ilist = dget(d_il, s.nloc);
  hn.numneigh = dget(d_nn, s.nloc);
  const std::vector<int> rows = dget(d_rows, (size_t)s.nloc * mem);
  hn.rows.assign(s.nloc, {});
  std::vector<int> flat;
  for (int i = 0; i < s.nloc; ++i) {
    hn.rows[i].assign(rows.begin() + (size_t)i * mem, rows.begin() + (size_t)i * mem + hn.numneigh[i]);
    flat.insert(flat.end(), hn.rows[i].begin(), hn.rows[i].end());
  }
  hn.make();
  wr(pre + "bn_numneigh.i32", hn.numneigh);
  wr(pre + "bn_rows.i32", flat);
  wr(pre + "bn_max.i32", &max_size, 1);
  // prod_env_mat_a_gpu with the device list made by convert_nlist_gpu_device
  const int width = std::max(256, (max_size + 255) / 256 * 256);
  int* gpu_mem = dnew<int>((size_t)s.nloc * width);
  deepmd::InputNlist gin;
  deepmd::convert_nlist_gpu_device(gin, hn.nl, gpu_mem, width);
  T *d_em = dnew<T>((size_t)s.nloc * s.nnei * 4), *d_ed = dnew<T>((size_t)s.nloc * s.nnei * 12),
    *d_rij = dnew<T>((size_t)s.nloc * s.nnei * 3);
  int* d_nl = dnew<int>((size_t)s.nloc * s.nnei);
  T *d_avg = dnew<T>(s.avg.size(), s.avg.data()), *d_std = dnew<T>(s.std_.size(), s.std_.data());
  deepmd::prod_env_mat_a_gpu<T>(d_em, d_ed, d_rij, d_nl, d_ext, d_ext_t, gin, nullptr, nullptr, width, d_avg, d_std,
                                s.nloc, nall, s.rcut, s.rcut_smth, s.sec);
  wr(pre + "em" + sfx, dget(d_em, (size_t)s.nloc * s.nnei * 4));
  wr(pre + "em_deriv" + sfx, dget(d_ed, (size_t)s.nloc * s.nnei * 12));
  wr(pre + "rij" + sfx, dget(d_rij, (size_t)s.nloc * s.nnei * 3));
  wr(pre + "nlist.i32", dget(d_nl, (size_t)s.nloc * s.nnei));
  int* d_fmt = dnew<int>((size_t)s.nloc * s.nnei);
  deepmd::format_nbor_list_gpu<T>(d_fmt, d_ext, d_ext_t, gin, nullptr, nullptr, width, s.nloc, nall, s.rcut, s.sec);
  wr(pre + "fmt_nlist.i32", dget(d_fmt, (size_t)s.nloc * s.nnei));
  deepmd::free_nlist_gpu_device(gin);
  for (void* p : {(void*)d_boxt, (void*)d_rec, (void*)d_coord, (void*)d_ci, (void*)d_type, (void*)d_ext,
                  (void*)d_ext_t, (void*)d_map, (void*)d_il, (void*)d_nn, (void*)d_rows, (void*)d_fn,
                  (void*)gpu_mem, (void*)d_em, (void*)d_ed, (void*)d_rij, (void*)d_nl, (void*)d_avg, (void*)d_std,
                  (void*)d_fmt})
    (void)hipFree(p);
}

int run_lib() {
  System<double> s = load_system<double>();
  std::vector<double> ext_c;
  std::vector<int> ext_t, mapping;
  HostNlist hn;
  geometry_chain<double>("", s, ext_c, ext_t, mapping, hn);
  const int nall = (int)ext_t.size(), nloc = s.nloc, nnei = s.nnei;
  // env_mat_nbor_update: the host list's pointers packed in a device mesh
  {
    std::vector<int> mesh(16, 0);
    mesh[0] = 0;
    mesh[1] = nloc;
    int* il = hn.ilist.data();
    int* nn = hn.numneigh.data();
    int** fn = hn.first.data();
    std::memcpy(&mesh[4], &il, sizeof(int*));
    std::memcpy(&mesh[8], &nn, sizeof(int*));
    std::memcpy(&mesh[12], &fn, sizeof(int**));
    int* d_mesh = dnew<int>(16, mesh.data());
    deepmd::InputNlist inlist, gpu_inlist;
    inlist.inum = nloc;  // the caller sets the row count; the mesh carries the pointers
    int max_nbor = 0;
    int* nbor_dev = nullptr;
    deepmd::env_mat_nbor_update(inlist, gpu_inlist, max_nbor, nbor_dev, d_mesh, 16);
    double *d_ext = dnew<double>(ext_c.size(), ext_c.data()), *d_avg = dnew<double>(s.avg.size(), s.avg.data()),
           *d_std = dnew<double>(s.std_.size(), s.std_.data());
    int* d_t = dnew<int>(nall, ext_t.data());
    double *d_em = dnew<double>((size_t)nloc * nnei * 4), *d_ed = dnew<double>((size_t)nloc * nnei * 12),
           *d_rij = dnew<double>((size_t)nloc * nnei * 3);
    int* d_nl = dnew<int>((size_t)nloc * nnei);
    deepmd::prod_env_mat_a_gpu<double>(d_em, d_ed, d_rij, d_nl, d_ext, d_t, gpu_inlist, nullptr, nullptr, max_nbor,
                                       d_avg, d_std, nloc, nall, s.rcut, s.rcut_smth, s.sec);
    wr("mesh_em.f64", dget(d_em, (size_t)nloc * nnei * 4));
    wr("mesh_max_nbor.i32", &max_nbor, 1);
    // use_nlist_map: the formatted list mapped onto the owners
    deepmd::use_nlist_map(d_nl, dnew<int>(nall, mapping.data()), nloc, nnei);
    wr("mapped_nlist.i32", dget(d_nl, (size_t)nloc * nnei));
    deepmd::free_nlist_gpu_device(gpu_inlist);
    (void)hipFree(nbor_dev);
  }
  // force / virial and their adjoints on the formatted list
  const std::vector<double> em_deriv = rd<double>("em_deriv.f64"), rij = rd<double>("rij.f64");
  const std::vector<int> nlist = rd<int>("nlist.i32");
  const std::vector<double> nd = rd<double>("nd.f64"), gf = rd<double>("gforce.f64"), gv = rd<double>("gvir.f64");
  double *d_nd = dnew<double>(nd.size(), nd.data()), *d_ed = dnew<double>(em_deriv.size(), em_deriv.data()),
         *d_rij = dnew<double>(rij.size(), rij.data());
  int* d_nl = dnew<int>(nlist.size(), nlist.data());
  double *d_f = dnew<double>(3 * (size_t)nall), *d_v = dnew<double>(9), *d_av = dnew<double>(9 * (size_t)nall);
  deepmd::prod_force_a_gpu<double>(d_f, d_nd, d_ed, d_nl, nloc, nall, nnei, 1);
  deepmd::prod_virial_a_gpu<double>(d_v, d_av, d_nd, d_ed, d_rij, d_nl, nloc, nall, nnei);
  wr("force.f64", dget(d_f, 3 * (size_t)nall));
  wr("virial.f64", dget(d_v, 9));
  wr("atom_virial.f64", dget(d_av, 9 * (size_t)nall));
  double *d_gf = dnew<double>(gf.size(), gf.data()), *d_gv = dnew<double>(9, gv.data()),
         *d_gn = dnew<double>(nd.size());
  deepmd::prod_force_grad_a_gpu<double>(d_gn, d_gf, d_ed, d_nl, nloc, nnei, 1);
  wr("gnet_f.f64", dget(d_gn, nd.size()));
  deepmd::prod_virial_grad_a_gpu<double>(d_gn, d_gv, d_ed, d_rij, d_nl, nloc, nnei);
  wr("gnet_v.f64", dget(d_gn, nd.size()));
  // tabulate on the reference's op-test data
  const std::vector<int> td = rd<int>("tab_dims.i32");
  const int tn = td[0], tk = td[1], M = td[2];
  const std::vector<double> tab = rd<double>("tab_table.f64"), info = rd<double>("tab_info.f64"),
                            ex = rd<double>("tab_emx.f64"), em = rd<double>("tab_em.f64"),
                            dy = rd<double>("tab_dy.f64");
  double *d_tab = dnew<double>(tab.size(), tab.data()), *d_ex = dnew<double>(ex.size(), ex.data()),
         *d_em = dnew<double>(em.size(), em.data()), *d_dy = dnew<double>(dy.size(), dy.data()),
         *d_out = dnew<double>((size_t)tn * 4 * M), *d_dx = dnew<double>(ex.size()), *d_dem = dnew<double>(em.size());
  deepmd::tabulate_fusion_se_a_gpu<double>(d_out, d_tab, info.data(), d_ex, d_em, nullptr, tn, tk, M);
  wr("tab_out.f64", dget(d_out, (size_t)tn * 4 * M));
  deepmd::tabulate_fusion_se_a_grad_gpu<double>(d_dx, d_dem, nullptr, d_tab, info.data(), d_ex, d_em, nullptr, d_dy,
                                                tn, tk, M);
  wr("tab_dx.f64", dget(d_dx, ex.size()));
  wr("tab_dem.f64", dget(d_dem, em.size()));
  deepmd::tabulate_fusion_se_a_grad_grad_gpu<double>(d_out, d_tab, info.data(), d_ex, d_em, nullptr, d_dx, d_dem,
                                                     nullptr, tn, tk, M);
  wr("tab_dz.f64", dget(d_out, (size_t)tn * 4 * M));

  // host-array entry points on the same inputs
  {
    deepmd::Region<double> hreg;
    deepmd::init_region_cpu(hreg, s.box.data());
    std::vector<double> c = s.coord;
    deepmd::normalize_coord_cpu<double>(c.data(), nloc, hreg);
    wr("cpu_norm.f64", c);
    const int mem_nall = 64 * nloc;
    std::vector<double> oc(3 * (size_t)mem_nall);
    std::vector<int> ot(mem_nall), om(mem_nall);
    int n2 = 0;
    if (deepmd::copy_coord_cpu<double>(oc.data(), ot.data(), om.data(), &n2, c.data(), s.atype.data(), nloc,
                                       mem_nall, s.rcut, hreg))
      throw std::runtime_error("copy_coord_cpu capacity");
    wr("cpu_ext_coord.f64", oc.data(), 3 * (size_t)n2);
    wr("cpu_mapping.i32", om.data(), n2);
    const int mem = 512;
    std::vector<int> il(nloc), nn(nloc), rows((size_t)nloc * mem);
    std::vector<int*> fp(nloc);
    for (int i = 0; i < nloc; ++i) fp[i] = rows.data() + (size_t)i * mem;
    deepmd::InputNlist hl(nloc, il.data(), nn.data(), fp.data());
    int mx = 0;
    if (deepmd::build_nlist_cpu<double>(hl, &mx, oc.data(), nloc, n2, mem, s.rcut, 1, ot.data()))
      throw std::runtime_error("build_nlist_cpu capacity");
    wr("cpu_bn_numneigh.i32", nn);
    std::vector<int> fmt((size_t)nloc * nnei);
    deepmd::format_nlist_cpu<double>(fmt.data(), hl, oc.data(), ot.data(), nloc, n2, s.rcut, s.sec);
    wr("cpu_fmt_nlist.i32", fmt);
    std::vector<double> e4((size_t)nloc * nnei * 4), e12((size_t)nloc * nnei * 12), r3((size_t)nloc * nnei * 3);
    std::vector<int> nl2((size_t)nloc * nnei);
    deepmd::prod_env_mat_a_cpu<double>(e4.data(), e12.data(), r3.data(), nl2.data(), oc.data(), ot.data(), hl, mx,
                                       s.avg.data(), s.std_.data(), nloc, n2, s.rcut, s.rcut_smth, s.sec);
    wr("cpu_em.f64", e4);
    wr("cpu_nlist.i32", nl2);
    std::vector<double> f(3 * (size_t)nall), v(9), av(9 * (size_t)nall), gn(nd.size());
    deepmd::prod_force_a_cpu<double>(f.data(), nd.data(), em_deriv.data(), nlist.data(), nloc, nall, nnei, 1);
    deepmd::prod_virial_a_cpu<double>(v.data(), av.data(), nd.data(), em_deriv.data(), rij.data(), nlist.data(), nloc,
                                      nall, nnei);
    wr("cpu_force.f64", f);
    wr("cpu_virial.f64", v);
    deepmd::prod_force_grad_a_cpu<double>(gn.data(), gf.data(), em_deriv.data(), nlist.data(), nloc, nnei, 1);
    wr("cpu_gnet_f.f64", gn);
    deepmd::prod_virial_grad_a_cpu<double>(gn.data(), gv.data(), em_deriv.data(), rij.data(), nlist.data(), nloc,
                                           nnei);
    wr("cpu_gnet_v.f64", gn);
    std::vector<double> out((size_t)tn * 4 * M), dx(ex.size()), dem(em.size());
    deepmd::tabulate_fusion_se_a_cpu<double>(out.data(), tab.data(), info.data(), ex.data(), em.data(), nullptr, tn,
                                             tk, M);
    wr("cpu_tab_out.f64", out);
    deepmd::tabulate_fusion_se_a_grad_cpu<double>(dx.data(), dem.data(), nullptr, tab.data(), info.data(), ex.data(),
                                                  em.data(), nullptr, dy.data(), tn, tk, M);
    wr("cpu_tab_dx.f64", dx);
    wr("cpu_tab_dem.f64", dem);
    deepmd::tabulate_fusion_se_a_grad_grad_cpu<double>(out.data(), tab.data(), info.data(), ex.data(), em.data(),
                                                       nullptr, dx.data(), dem.data(), nullptr, tn, tk, M);
    wr("cpu_tab_dz.f64", out);
  }
  // float instantiations of the geometry and environment chain
  {
    System<float> sf = load_system<float>();
    std::vector<float> ec;
    std::vector<int> et, mp;
    HostNlist hf;
    geometry_chain<float>("f32_", sf, ec, et, mp, hf);
  }
  std::cout << "lib ok nall=" << nall << std::endl;
  return 0;
}

deepmd::InputNlist rows_nlist(HostNlist& h, const std::string& pre) {
  h.ilist = rd<int>(pre + "ilist.i32");
  h.numneigh = rd<int>(pre + "numneigh.i32");
  const std::vector<int> flat = rd<int>(pre + "rows.i32");
  h.rows.clear();
  size_t off = 0;
  for (int n : h.numneigh) {
    h.rows.emplace_back(flat.begin() + off, flat.begin() + off + n);
    off += n;
  }
  h.make();
  return h.nl;
}

int run_deeppot() {
  const std::vector<int> prec = rd<int>("dp_precision.i32");
  dpmi::DeepPot dp(g_dir + "/model.dpmi", 0, prec[0] ? "mixed" : "fp64");
  std::string tm;
  dp.get_type_map(tm);
  const double rc = dp.cutoff();
  const int nt = dp.numb_types();
  wr("dp_cutoff.f64", &rc, 1);
  wr("dp_ntypes.i32", &nt, 1);
  wr("dp_type_map.i32", std::vector<int>(tm.begin(), tm.end()));
  // periodic, no list: double and float
  const std::vector<double> c = rd<double>("pbc_coord.f64"), box = rd<double>("pbc_box.f64");
  const std::vector<int> t = rd<int>("pbc_atype.i32");
  double e = 0;
  std::vector<double> f, v, ae, av;
  dp.compute(e, f, v, ae, av, c, t, box);
  wr("pbc_e.f64", &e, 1);
  wr("pbc_f.f64", f);
  wr("pbc_v.f64", v);
  wr("pbc_ae.f64", ae);
  wr("pbc_av.f64", av);
  std::vector<float> cf(c.begin(), c.end()), bf(box.begin(), box.end()), ff, vf;
  double ef = 0;
  dp.compute(ef, ff, vf, cf, t, bf);
  wr("pbcf_e.f64", &ef, 1);
  wr("pbcf_f.f32", ff);
  // no box
  const std::vector<double> c0 = rd<double>("nopbc_coord.f64");
  const std::vector<int> t0 = rd<int>("nopbc_atype.i32");
  dp.compute(e, f, v, ae, av, c0, t0, std::vector<double>());
  wr("nopbc_e.f64", &e, 1);
  wr("nopbc_f.f64", f);
  wr("nopbc_v.f64", v);
  wr("nopbc_ae.f64", ae);
  wr("nopbc_av.f64", av);
  // LAMMPS style: caller's ghosts and list; ago 0, then ago 1 on moved atoms
  const std::vector<double> lc = rd<double>("lmp_coord.f64"), lc2 = rd<double>("lmp_coord2.f64");
  const std::vector<int> lt = rd<int>("lmp_atype.i32"), ng = rd<int>("lmp_nghost.i32");
  HostNlist h;
  deepmd::InputNlist nl = rows_nlist(h, "lmp_");
  dp.compute(e, f, v, ae, av, lc, lt, box, ng[0], nl, 0);
  wr("lmp_e.f64", &e, 1);
  wr("lmp_f.f64", f);
  wr("lmp_v.f64", v);
  wr("lmp_ae.f64", ae);
  wr("lmp_av.f64", av);
  dp.compute(e, f, v, lc2, lt, box, ng[0], nl, 1);
  wr("lmp2_e.f64", &e, 1);
  wr("lmp2_f.f64", f);
  std::cout << "deeppot ok" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: dpmi_harness lib|deeppot <dir>" << std::endl;
    return 2;
  }
  g_dir = argv[2];
  try {
    const std::string mode = argv[1];
    if (mode == "lib") return run_lib();
    if (mode == "deeppot") return run_deeppot();
    std::cerr << "unknown mode " << mode << std::endl;
    return 2;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
