// dpmi model files: an se_e2_a energy model as plain arrays (descriptor
// settings, davg/dstd, embedding and fitting weights, bias_atom_e and the
// compression settings), loaded natively.  The compressed tables are built on
// the GPU from the embedding weights (tablegen.hip), so a model is ready for
// inference without Python.  Layout (little-endian; the writer is
// dpmi.model.save_model_file):
//   char[8] "DPMISEA1"
//   int32 ntypes, type_one_side, axis_neuron, n_embed_layers, n_fit_layers
//   float64 rcut, rcut_smth
//   int32 sel[ntypes], embed_width[n_embed_layers + 1] (1, ..., M),
//         fit_width[n_fit_layers + 1] (M*A, ..., 1),
//         embed_has_idt[n_embed_layers], fit_flags[n_fit_layers]
//         (bit 0 tanh, bit 1 residual, bit 2 has idt)
//   float64 davg[ntypes * nnei * 4], dstd[ntypes * nnei * 4]
//   per embedding net (ntypes or ntypes^2, net = tj * ntypes + ti), per layer:
//     w [n_in * n_out], b [n_out], idt [n_out] if present
//   per type, per fitting layer: the same
//   float64 bias_atom_e[ntypes]
//   float64 min_nbor_dist, extrapolate, stride0, stride1, check_frequency
//   int32 ntype_names, then per name int32 length + bytes
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <iterator>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dpmi_common.h"
#include "engine.h"
#include "kernels.h"
#include "model_file.h"

namespace dpmi {

namespace {

struct Reader {
  std::vector<char> buf;
  size_t pos = 0;
  void need(size_t n) const {
    if (pos + n > buf.size()) throw std::invalid_argument("model file is truncated");
  }
  template <typename T>
  T one() {
    need(sizeof(T));
    T v;
    std::memcpy(&v, buf.data() + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  template <typename T>
  std::vector<T> many(size_t n) {
    need(n * sizeof(T));
    std::vector<T> v(n);
    if (n) std::memcpy(v.data(), buf.data() + pos, n * sizeof(T));
    pos += n * sizeof(T);
    return v;
  }
};

double spline5_switch(double x, double rmin, double rmax) {
#pragma clang fp contract(off)
  // tabulate.py:488-496
  if (x < rmin) return 1.0;
  if (x < rmax) {
    const double u = (x - rmin) / (rmax - rmin);
    return u * u * u * (-6 * u * u + 15 * u - 10) + 1;
  }
  return 0.0;
}

}  // namespace

ModelFile read_model_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::invalid_argument("cannot open model file " + path);
  Reader r;
  r.buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  r.need(8);
  if (std::memcmp(r.buf.data(), "DPMISEA1", 8) != 0) throw std::invalid_argument(path + " is not a dpmi model file");
  r.pos = 8;
  ModelFile m;
  m.ntypes = r.one<int>();
  m.type_one_side = r.one<int>();
  m.axis = r.one<int>();
  const int ne = r.one<int>(), nf = r.one<int>();
  if (m.ntypes <= 0 || m.ntypes > 127 || ne <= 0 || nf <= 1 || ne > 64 || nf > 64)
    throw std::invalid_argument("model file header is out of range");
  m.rcut = r.one<double>();
  m.rcut_smth = r.one<double>();
  m.sel = r.many<int>(m.ntypes);
  m.embed_width = r.many<int>(ne + 1);
  m.fit_width = r.many<int>(nf + 1);
  const std::vector<int> e_idt = r.many<int>(ne), f_flags = r.many<int>(nf);
  int nnei = 0;
  for (int s : m.sel) {
    if (s < 0) throw std::invalid_argument("negative sel");
    nnei += s;
  }
  for (int w : m.embed_width)
    if (w <= 0) throw std::invalid_argument("embedding widths must be positive");
  for (int w : m.fit_width)
    if (w <= 0) throw std::invalid_argument("fitting widths must be positive");
  m.davg = r.many<double>((size_t)m.ntypes * nnei * 4);
  m.dstd = r.many<double>((size_t)m.ntypes * nnei * 4);
  const int nnets = m.type_one_side ? m.ntypes : m.ntypes * m.ntypes;
  auto read_net = [&](const std::vector<int>& width, int nl, auto has_idt) {
    ModelFile::Net net;
    for (int l = 0; l < nl; ++l) {
      const size_t ni = width[l], no = width[l + 1];
      net.w.push_back(r.many<double>(ni * no));
      net.b.push_back(r.many<double>(no));
      net.idt.push_back(has_idt(l) ? r.many<double>(no) : std::vector<double>());
    }
    return net;
  };
  for (int k = 0; k < nnets; ++k) m.embed.push_back(read_net(m.embed_width, ne, [&](int l) { return e_idt[l] != 0; }));
  for (int t = 0; t < m.ntypes; ++t)
    m.fit.push_back(read_net(m.fit_width, nf, [&](int l) { return (f_flags[l] & 4) != 0; }));
  for (int l = 0; l < nf; ++l) {
    m.fit_act.push_back(f_flags[l] & 1);
    m.fit_resnet.push_back((f_flags[l] >> 1) & 1);
  }
  m.bias_atom_e = r.many<double>(m.ntypes);
  m.min_nbor_dist = r.one<double>();
  m.extrapolate = r.one<double>();
  m.stride0 = r.one<double>();
  m.stride1 = r.one<double>();
  m.check_frequency = r.one<double>();
  const int nnames = r.one<int>();
  if (nnames < 0 || nnames > 4096) throw std::invalid_argument("type map size is out of range");
  for (int k = 0; k < nnames; ++k) {
    const int len = r.one<int>();
    if (len < 0 || len > 4096) throw std::invalid_argument("type name length is out of range");
    const std::vector<char> c = r.many<char>(len);
    m.type_map.emplace_back(c.begin(), c.end());
  }
  return m;
}

// tabulate.py:67-140, 468-486: per centre type floor(-davg/dstd) and
// ceil((s(d_min)/d_min - davg)/dstd) of the first environment column; nets of
// a one-sided model take the extremes over the types with sel > 0.
std::vector<std::pair<double, double>> model_table_ranges(const ModelFile& m) {
#pragma clang fp contract(off)
  const int nt = m.ntypes;
  int nnei = 0;
  for (int s : m.sel) nnei += s;
  const double sw = spline5_switch(m.min_nbor_dist, m.rcut_smth, m.rcut);
  std::vector<double> lo(nt), up(nt);
  for (int t = 0; t < nt; ++t) {
    const double a = m.davg[(size_t)t * nnei * 4], s = m.dstd[(size_t)t * nnei * 4];
    lo[t] = std::floor(-a / s);
    up[t] = std::ceil(((1.0 / m.min_nbor_dist) * sw - a) / s);
  }
  const int nnets = m.type_one_side ? nt : nt * nt;
  std::vector<std::pair<double, double>> out;
  for (int k = 0; k < nnets; ++k) {
    if (m.type_one_side) {
      double l = INFINITY, u = -INFINITY;
      for (int t = 0; t < nt; ++t)
        if (m.sel[t] > 0) {
          l = std::min(l, lo[t]);
          u = std::max(u, up[t]);
        }
      out.emplace_back(l, u);
    } else {
      out.emplace_back(lo[k % nt], up[k % nt]);
    }
  }
  return out;
}

SeaModel* build_model(const ModelFile& m, hipStream_t st) {
  const int nt = m.ntypes, nnets = (int)m.embed.size();
  const int ne = (int)m.embed_width.size() - 1, nf = (int)m.fit_width.size() - 1;
  const auto ranges = model_table_ranges(m);
  std::vector<std::vector<double>> tables(nnets);
  std::vector<int> nspline(nnets);
  std::vector<double> infos;
  for (int k = 0; k < nnets; ++k) {
    TabNet net;
    net.width = m.embed_width;
    for (int l = 0; l < ne; ++l) {
      net.w.push_back(m.embed[k].w[l].data());
      net.b.push_back(m.embed[k].b[l].data());
      net.idt.push_back(m.embed[k].idt[l].empty() ? nullptr : m.embed[k].idt[l].data());
    }
    const double ll = ranges[k].first, uu = ranges[k].second;
    build_table_device(net, ll, uu, m.extrapolate, m.stride0, m.stride1, tables[k], nspline[k], st);
    for (double v : {ll, uu, uu * m.extrapolate, m.stride0, m.stride1, m.check_frequency}) infos.push_back(v);
  }
  std::vector<const double*> tp(nnets), fw, fb, fi;
  for (int k = 0; k < nnets; ++k) tp[k] = tables[k].data();
  for (int t = 0; t < nt; ++t)
    for (int l = 0; l < nf; ++l) {
      fw.push_back(m.fit[t].w[l].data());
      fb.push_back(m.fit[t].b[l].data());
      fi.push_back(m.fit[t].idt[l].empty() ? nullptr : m.fit[t].idt[l].data());
    }
  dpmi_sea_model_desc d{};
  d.ntypes = nt;
  d.sel = m.sel.data();
  d.rcut = (float)m.rcut;
  d.rcut_smth = (float)m.rcut_smth;
  d.last_layer_size = m.embed_width.back();
  d.axis_neuron = m.axis;
  d.type_one_side = m.type_one_side;
  d.davg = m.davg.data();
  d.dstd = m.dstd.data();
  d.nnets = nnets;
  d.nspline = nspline.data();
  d.tables = tp.data();
  d.table_info = infos.data();
  d.fit_nlayers = nf;
  d.fit_width = m.fit_width.data();
  d.fit_w = fw.data();
  d.fit_b = fb.data();
  d.fit_idt = fi.data();
  d.fit_act = m.fit_act.data();
  d.fit_resnet = m.fit_resnet.data();
  d.bias_atom_e = m.bias_atom_e.data();
  auto* sm = new SeaModel(d);
  sm->type_map = m.type_map;
  return sm;
}

}  // namespace dpmi
