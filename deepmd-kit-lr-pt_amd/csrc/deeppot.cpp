// dpmi::DeepPot (include/dpmi_deeppot.h): host vectors in and out around the
// dpmi_sea_* engine, as DeepPotPT::compute does around its TorchScript module
// (source/api_cc/src/DeepPotPT.cc:171-500).
#include "../../include/dpmi_deeppot.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/dpmi.h"
#include "dpmi_common.h"
#include "engine.h"
#include "kernels.h"

namespace dpmi {

namespace {

void ok_or_throw(int rc) {
  if (rc == DPMI_OK) return;
  if (rc == DPMI_ERR_NLIST_CAPACITY) throw deepmd::deepmd_exception_nlist_capacity(dpmi_last_error());
  if (rc == DPMI_ERR_MEMORY) throw deepmd::deepmd_exception_oom(dpmi_last_error());
  throw deepmd::deepmd_exception(dpmi_last_error());
}
void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw deepmd::deepmd_exception(std::string(what) + ": " + hipGetErrorString(e));
}

// grow-only device buffer
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  T* get(size_t count) {
    if (count > n) {
      if (p) (void)hipFree(p);
      p = nullptr;
      hip_ok(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
      n = count;
    }
    return p;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename V>
std::vector<double> to_double(const std::vector<V>& v) {
  return std::vector<double>(v.begin(), v.end());
}
template <typename V>
void from_double(std::vector<V>& out, const std::vector<double>& v) {
  out.assign(v.begin(), v.end());
}

}  // namespace

struct DeepPot::Impl {
  dpmi_sea_model* model = nullptr;
  int precision = DPMI_FP64;
  int gpu = 0;
  hipStream_t st = nullptr;
  double rcut = 0;
  int ntypes = 0;
  std::string type_map;
  // engines keyed by the atom types they were created for
  dpmi_sea_engine* eng_pbc = nullptr;
  std::vector<int> key_pbc;
  dpmi_sea_engine* eng_low = nullptr;
  std::vector<int> key_low;
  DBuf<double> coord, force, virial, energy, atom_e, atom_v;
  DBuf<int> atype, ilist, numneigh, jlist;
  int jstride = 0, rows_uploaded = 0;

  ~Impl() {
    if (eng_pbc) dpmi_sea_engine_destroy(eng_pbc);
    if (eng_low) dpmi_sea_engine_destroy(eng_low);
    if (model) dpmi_sea_model_destroy(model);
    if (st) (void)hipStreamDestroy(st);
  }

  dpmi_sea_engine* engine(dpmi_sea_engine*& e, std::vector<int>& key, const std::vector<int>& types) {
    if (e && key == types) return e;
    if (e) dpmi_sea_engine_destroy(e);
    e = nullptr;
    const int n = (int)types.size();
    ok_or_throw(dpmi_sea_engine_create(&e, model, 1, &n, types.data(), precision, 0.0));
    key = types;
    return e;
  }

  // engine step on the engine's own ghosts (periodic box)
  void run_pbc(double& ener, std::vector<double>& f, std::vector<double>& v, std::vector<double>* ae,
               std::vector<double>* av, const std::vector<double>& c, const std::vector<int>& t,
               const std::vector<double>& box) {
    const size_t n = t.size();
    dpmi_sea_engine* e = engine(eng_pbc, key_pbc, t);
    double* dc = coord.get(3 * n);
    hip_ok(hipMemcpyAsync(dc, c.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, st), "copy coord");
    ok_or_throw(dpmi_sea_engine_compute(e, dc, box.data(), energy.get(1), force.get(3 * n), virial.get(9),
                                        ae ? atom_e.get(n) : nullptr, av ? atom_v.get(9 * n) : nullptr, st));
    ok_or_throw(dpmi_sea_engine_status(e));
    download(ener, f, v, ae, av, n, n, n);
  }

  // lower path: nall caller atoms, rows already on the device
  void run_lower(double& ener, std::vector<double>& f, std::vector<double>& v, std::vector<double>* ae,
                 std::vector<double>* av, const std::vector<double>& c, const std::vector<int>& t, int nloc) {
    const size_t nall = t.size();
    dpmi_sea_engine* e = engine(eng_low, key_low, std::vector<int>(t.begin(), t.begin() + nloc));
    double* dc = coord.get(3 * nall);
    int* dt = atype.get(nall);
    hip_ok(hipMemcpyAsync(dc, c.data(), sizeof(double) * 3 * nall, hipMemcpyHostToDevice, st), "copy coord");
    hip_ok(hipMemcpyAsync(dt, t.data(), sizeof(int) * nall, hipMemcpyHostToDevice, st), "copy atype");
    ok_or_throw(dpmi_sea_engine_compute_lower(e, dc, dt, (int)nall, nloc, ilist.p, numneigh.p, jlist.p, jstride,
                                              energy.get(1), force.get(3 * nall), virial.get(9),
                                              atom_e.get(std::max(nloc, 1)), av ? atom_v.get(9 * nall) : nullptr,
                                              st));
    ok_or_throw(dpmi_sea_engine_status(e));
    download(ener, f, v, ae, av, nall, nloc, nall);
  }

  void download(double& ener, std::vector<double>& f, std::vector<double>& v, std::vector<double>* ae,
                std::vector<double>* av, size_t nf, size_t ne, size_t nv) {
    f.resize(3 * nf);
    v.resize(9);
    hip_ok(hipMemcpyAsync(&ener, energy.p, sizeof(double), hipMemcpyDeviceToHost, st), "copy energy");
    hip_ok(hipMemcpyAsync(f.data(), force.p, sizeof(double) * f.size(), hipMemcpyDeviceToHost, st), "copy force");
    hip_ok(hipMemcpyAsync(v.data(), virial.p, sizeof(double) * 9, hipMemcpyDeviceToHost, st), "copy virial");
    if (ae) {
      ae->assign(nf, 0.0);  // ghosts (nf > ne) carry no energy
      hip_ok(hipMemcpyAsync(ae->data(), atom_e.p, sizeof(double) * ne, hipMemcpyDeviceToHost, st), "copy atom_e");
    }
    if (av) {
      av->resize(9 * nv);
      hip_ok(hipMemcpyAsync(av->data(), atom_v.p, sizeof(double) * 9 * nv, hipMemcpyDeviceToHost, st), "copy av");
    }
    hip_ok(hipStreamSynchronize(st), "synchronize");
  }

  // rows of the caller's list on the device, width jstride (a multiple of 64)
  void upload_rows(const deepmd::InputNlist& l) {
    int mx = 0;
    for (int i = 0; i < l.inum; ++i) mx = std::max(mx, l.numneigh[i]);
    if (mx > 4096) throw deepmd::deepmd_exception_nlist_capacity("a neighbour row holds more than 4096 atoms");
    jstride = std::max(64, (mx + 63) / 64 * 64);
    std::vector<int> packed((size_t)l.inum * jstride, -1);
    // common.cc:297-316: indices carry the caller's mask (LAMMPS special bits)
    for (int i = 0; i < l.inum; ++i)
      for (int k = 0; k < l.numneigh[i]; ++k) packed[(size_t)i * jstride + k] = l.firstneigh[i][k] & l.mask;
    hip_ok(hipMemcpyAsync(ilist.get(l.inum), l.ilist, sizeof(int) * l.inum, hipMemcpyHostToDevice, st), "ilist");
    hip_ok(hipMemcpyAsync(numneigh.get(l.inum), l.numneigh, sizeof(int) * l.inum, hipMemcpyHostToDevice, st),
           "numneigh");
    hip_ok(hipMemcpyAsync(jlist.get(packed.size()), packed.data(), sizeof(int) * packed.size(), hipMemcpyHostToDevice,
                          st),
           "jlist");
    hip_ok(hipStreamSynchronize(st), "synchronize");
    rows_uploaded = l.inum;
  }

  // no periodic images: every pair of local atoms within rcut is a candidate
  void rows_without_box(const std::vector<double>& c, const std::vector<int>& t) {
    const int n = (int)t.size();
    double* dc = coord.get(3 * (size_t)n);
    int* dt = atype.get(n);
    hip_ok(hipMemcpyAsync(dc, c.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, st), "copy coord");
    hip_ok(hipMemcpyAsync(dt, t.data(), sizeof(int) * n, hipMemcpyHostToDevice, st), "copy atype");
    const double rc2 = (double)((float)rcut * (float)rcut);
    for (int cap = 256;; cap *= 2) {
      cap = std::min(cap, 4096);
      launch_brute_nlist<double>(dc, n, n, cap, rc2, 1, dt, ilist.get(n), numneigh.get(n),
                                 jlist.get((size_t)n * cap), st);
      std::vector<int> nn(n);
      hip_ok(hipMemcpyAsync(nn.data(), numneigh.p, sizeof(int) * n, hipMemcpyDeviceToHost, st), "numneigh");
      hip_ok(hipStreamSynchronize(st), "synchronize");
      const int mx = n ? *std::max_element(nn.begin(), nn.end()) : 0;
      if (mx <= cap) {
        jstride = cap;
        break;
      }
      if (cap == 4096) throw deepmd::deepmd_exception_nlist_capacity("more than 4096 atoms within rcut of one atom");
    }
    rows_uploaded = 0;  // the next list call uploads its own rows
  }
};

DeepPot::DeepPot() : impl_(nullptr) {}
DeepPot::DeepPot(const std::string& model, const int& gpu_rank, const std::string& precision) {
  init(model, gpu_rank, precision);
}
DeepPot::~DeepPot() { delete impl_; }

void DeepPot::init(const std::string& model, const int& gpu_rank, const std::string& precision) {
  if (impl_) return;  // the reference ignores a second init
  auto* im = new Impl();
  try {
    if (precision != "fp64" && precision != "mixed") throw deepmd::deepmd_exception("precision is fp64 or mixed");
    im->precision = precision == "fp64" ? DPMI_FP64 : DPMI_MIXED;
    int ndev = 0;
    hip_ok(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (ndev <= 0) throw deepmd::deepmd_exception("no HIP device: dpmi runs on the GPU only");
    im->gpu = gpu_rank % ndev;
    hip_ok(hipSetDevice(im->gpu), "hipSetDevice");
    hip_ok(hipStreamCreate(&im->st), "hipStreamCreate");
    ok_or_throw(dpmi_sea_model_load(&im->model, model.c_str(), im->st));
    char names[4096];
    ok_or_throw(dpmi_sea_model_info(im->model, &im->ntypes, &im->rcut, nullptr, names, sizeof(names)));
    im->type_map = names;
  } catch (...) {
    delete im;
    throw;
  }
  impl_ = im;
}

double DeepPot::cutoff() const { return impl_ ? impl_->rcut : 0.0; }
int DeepPot::numb_types() const { return impl_ ? impl_->ntypes : 0; }
void DeepPot::get_type_map(std::string& type_map) { type_map = impl_ ? impl_->type_map : std::string(); }

namespace {
template <typename V>
void no_params(const std::vector<V>& fparam, const std::vector<V>& aparam) {
  if (!fparam.empty() || !aparam.empty())
    throw deepmd::deepmd_exception("this se_e2_a energy model takes no frame or atomic parameters");
}
}  // namespace

template <typename VALUETYPE>
void DeepPot::compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
                      std::vector<VALUETYPE>& atom_energy, std::vector<VALUETYPE>& atom_virial,
                      const std::vector<VALUETYPE>& coord, const std::vector<int>& atype,
                      const std::vector<VALUETYPE>& box, const std::vector<VALUETYPE>& fparam,
                      const std::vector<VALUETYPE>& aparam) {
  if (!impl_) throw deepmd::deepmd_exception("DeepPot is not initialised");
  no_params(fparam, aparam);
  hip_ok(hipSetDevice(impl_->gpu), "hipSetDevice");
  const std::vector<double> c = to_double(coord);
  if (c.size() != 3 * atype.size()) throw deepmd::deepmd_exception("coord must hold 3 values per atom");
  std::vector<double> f, v, ae, av;
  if (box.size() == 9) {
    impl_->run_pbc(ener, f, v, &ae, &av, c, atype, to_double(box));
  } else if (box.empty()) {
    impl_->rows_without_box(c, atype);
    impl_->run_lower(ener, f, v, &ae, &av, c, atype, (int)atype.size());
  } else {
    throw deepmd::deepmd_exception("box must hold 9 values or none");
  }
  from_double(force, f);
  from_double(virial, v);
  from_double(atom_energy, ae);
  from_double(atom_virial, av);
}

template <typename VALUETYPE>
void DeepPot::compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
                      const std::vector<VALUETYPE>& coord, const std::vector<int>& atype,
                      const std::vector<VALUETYPE>& box, const std::vector<VALUETYPE>& fparam,
                      const std::vector<VALUETYPE>& aparam) {
  if (!impl_) throw deepmd::deepmd_exception("DeepPot is not initialised");
  no_params(fparam, aparam);
  hip_ok(hipSetDevice(impl_->gpu), "hipSetDevice");
  const std::vector<double> c = to_double(coord);
  if (c.size() != 3 * atype.size()) throw deepmd::deepmd_exception("coord must hold 3 values per atom");
  std::vector<double> f, v;
  if (box.size() == 9) {
    impl_->run_pbc(ener, f, v, nullptr, nullptr, c, atype, to_double(box));
  } else if (box.empty()) {
    impl_->rows_without_box(c, atype);
    impl_->run_lower(ener, f, v, nullptr, nullptr, c, atype, (int)atype.size());
  } else {
    throw deepmd::deepmd_exception("box must hold 9 values or none");
  }
  from_double(force, f);
  from_double(virial, v);
}

template <typename VALUETYPE>
void DeepPot::compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
                      std::vector<VALUETYPE>& atom_energy, std::vector<VALUETYPE>& atom_virial,
                      const std::vector<VALUETYPE>& coord, const std::vector<int>& atype,
                      const std::vector<VALUETYPE>& box, const int nghost, const deepmd::InputNlist& inlist,
                      const int& ago, const std::vector<VALUETYPE>& fparam, const std::vector<VALUETYPE>& aparam) {
  if (!impl_) throw deepmd::deepmd_exception("DeepPot is not initialised");
  no_params(fparam, aparam);
  hip_ok(hipSetDevice(impl_->gpu), "hipSetDevice");
  const int nall = (int)atype.size(), nloc = nall - nghost;
  if (nghost < 0 || nloc < 0) throw deepmd::deepmd_exception("nghost must be in [0, natoms]");
  if (inlist.inum != nloc) throw deepmd::deepmd_exception("the neighbour list must hold one row per local atom");
  const std::vector<double> c = to_double(coord);
  if (c.size() != 3 * (size_t)nall) throw deepmd::deepmd_exception("coord must hold 3 values per atom");
  if (ago == 0 || impl_->rows_uploaded != nloc) impl_->upload_rows(inlist);
  std::vector<double> f, v, ae, av;
  impl_->run_lower(ener, f, v, &ae, &av, c, atype, nloc);
  from_double(force, f);
  from_double(virial, v);
  from_double(atom_energy, ae);
  from_double(atom_virial, av);
  (void)box;  // the caller's ghosts already carry the periodic images
}

template <typename VALUETYPE>
void DeepPot::compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
                      const std::vector<VALUETYPE>& coord, const std::vector<int>& atype,
                      const std::vector<VALUETYPE>& box, const int nghost, const deepmd::InputNlist& inlist,
                      const int& ago, const std::vector<VALUETYPE>& fparam, const std::vector<VALUETYPE>& aparam) {
  std::vector<VALUETYPE> ae, av;
  compute(ener, force, virial, ae, av, coord, atype, box, nghost, inlist, ago, fparam, aparam);
}

#define DPMI_DEEPPOT_INST(V)                                                                                       \
  template void DeepPot::compute<V>(double&, std::vector<V>&, std::vector<V>&, const std::vector<V>&,              \
                                    const std::vector<int>&, const std::vector<V>&, const std::vector<V>&,         \
                                    const std::vector<V>&);                                                        \
  template void DeepPot::compute<V>(double&, std::vector<V>&, std::vector<V>&, std::vector<V>&, std::vector<V>&,   \
                                    const std::vector<V>&, const std::vector<int>&, const std::vector<V>&,         \
                                    const std::vector<V>&, const std::vector<V>&);                                 \
  template void DeepPot::compute<V>(double&, std::vector<V>&, std::vector<V>&, const std::vector<V>&,              \
                                    const std::vector<int>&, const std::vector<V>&, const int,                     \
                                    const deepmd::InputNlist&, const int&, const std::vector<V>&,                  \
                                    const std::vector<V>&);                                                        \
  template void DeepPot::compute<V>(double&, std::vector<V>&, std::vector<V>&, std::vector<V>&, std::vector<V>&,   \
                                    const std::vector<V>&, const std::vector<int>&, const std::vector<V>&,         \
                                    const int, const deepmd::InputNlist&, const int&, const std::vector<V>&,       \
                                    const std::vector<V>&);
DPMI_DEEPPOT_INST(double)
DPMI_DEEPPOT_INST(float)

}  // namespace dpmi
