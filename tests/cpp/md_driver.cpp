// A small LAMMPS stand-in that runs NVE molecular dynamics through the pair
// style dpmi::lmp::PairDeepMD (include/dpmi_pair_deepmd.h), used by
// tests/test_gpu_lmp_md.py.  It keeps LAMMPS's single-process data flow:
//   * restricted triclinic domain (boxlo, h = xx yy zz yz xz xy), atoms
//     remapped into the box at each re-neighbouring;
//   * ghost atoms: every periodic image within cutghost = cutoff + skin of the
//     box (in lamda coordinates, as Comm does for triclinic boxes), updated
//     from their owners between re-neighbourings (forward communication);
//   * a full neighbour list of the local atoms over local and ghost atoms
//     within cutoff + skin, rows in arbitrary order (optionally shuffled, and
//     with special-bond bits set above NEIGHMASK); Neighbor::decide semantics:
//     ago counts the steps since the last build, rebuild when an atom has moved
//     more than skin / 2;
//   * newton on: ghost forces (and per-atom centroid virials) are summed onto
//     their owners after the pair compute (reverse communication);
//   * velocity Verlet (fix nve) in metal units.
//
//   dpmi_md <dir> <model> <precision> <steps> <dt ps> <skin> <dump_every>
//           [names <n1> <n2> ...] [shuffle] [special] [atom] [errors]
// Inputs in <dir>: box.f64 (rows a, b, c of a lower-triangular cell),
// boxlo.f64 (3), coord.f64, vel.f64 (A/ps), type.i32 (1-based), mass.f64
// (per LAMMPS type).  Outputs: thermo.f64 (per step: step, pe, ke, etot, ago,
// builds, virial xx yy zz xy xz yz), and at every dump step s:
// dump<s>_x.f64, dump<s>_f.f64 (owner-folded), with "atom" also
// dump<s>_eatom.f64 and dump<s>_cvatom.f64 (owner-folded, LAMMPS order).
// "errors" only runs the pair style's error contract and writes errors.i32.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "dpmi_pair_deepmd.h"

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
void wr(const std::string& name, const std::vector<T>& v) {
  std::ofstream f(g_dir + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

constexpr double kMvv2e = 1.0364269e-4;  // metal units: g/mol A^2/ps^2 -> eV
constexpr double kFtm2v = 1.0 / 1.0364269e-4;

struct Domain {
  double boxlo[3];
  double h[6];      // xx yy zz yz xz xy
  double h_inv[6];  // as LAMMPS Domain::set_global_box
  void set(const std::vector<double>& box, const std::vector<double>& lo) {
    for (int d = 0; d < 3; ++d) boxlo[d] = lo[d];
    h[0] = box[0];
    h[1] = box[4];
    h[2] = box[8];
    h[3] = box[7];
    h[4] = box[6];
    h[5] = box[3];
    h_inv[0] = 1.0 / h[0];
    h_inv[1] = 1.0 / h[1];
    h_inv[2] = 1.0 / h[2];
    h_inv[3] = -h[3] / (h[1] * h[2]);
    h_inv[4] = (h[3] * h[5] - h[1] * h[4]) / (h[0] * h[1] * h[2]);
    h_inv[5] = -h[5] / (h[0] * h[1]);
  }
  void x2lamda(const double* x, double* l) const {
    const double dx = x[0] - boxlo[0], dy = x[1] - boxlo[1], dz = x[2] - boxlo[2];
    l[0] = h_inv[0] * dx + h_inv[5] * dy + h_inv[4] * dz;
    l[1] = h_inv[1] * dy + h_inv[3] * dz;
    l[2] = h_inv[2] * dz;
  }
  void lamda2x(const double* l, double* x) const {
    x[0] = h[0] * l[0] + h[5] * l[1] + h[4] * l[2] + boxlo[0];
    x[1] = h[1] * l[1] + h[3] * l[2] + boxlo[1];
    x[2] = h[2] * l[2] + boxlo[2];
  }
  // image shift n (in cell vectors) as a Cartesian vector
  void shift(const int* n, double* s) const {
    s[0] = h[0] * n[0] + h[5] * n[1] + h[4] * n[2];
    s[1] = h[1] * n[1] + h[3] * n[2];
    s[2] = h[2] * n[2];
  }
  // cutghost in lamda units per dimension (1 / face distance per unit length)
  void cut_lamda(double cut, double* c) const {
    c[0] = cut * std::sqrt(h_inv[0] * h_inv[0] + h_inv[5] * h_inv[5] + h_inv[4] * h_inv[4]);
    c[1] = cut * std::sqrt(h_inv[1] * h_inv[1] + h_inv[3] * h_inv[3]);
    c[2] = cut * h_inv[2];
  }
};

struct Md {
  Domain dom;
  int nlocal = 0, nghost = 0, ntypes = 0;
  std::vector<std::array<double, 3>> x, v, f, xhold;
  std::vector<int> type;
  std::vector<double> mass;
  std::vector<int> owner;                 // per ghost (index - nlocal)
  std::vector<std::array<int, 3>> image;  // per ghost
  // neighbour list storage
  std::vector<int> ilist, numneigh;
  std::vector<std::vector<int>> rows;
  std::vector<int*> first;
  std::vector<double*> xp, fp;
  int ago = 0, builds = 0;
  double cutneigh = 0.0, skin = 0.0;
  bool shuffle = false, special = false;
  unsigned long long lcg = 12345;

  unsigned next_rand() {
    lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
    return (unsigned)(lcg >> 33);
  }

  void pbc() {
    for (int i = 0; i < nlocal; ++i) {
      double l[3];
      dom.x2lamda(x[i].data(), l);
      for (int d = 0; d < 3; ++d) {
        l[d] -= std::floor(l[d]);
        if (l[d] >= 1.0) l[d] = 0.0;
      }
      dom.lamda2x(l, x[i].data());
    }
  }

  void borders() {
    double c[3];
    dom.cut_lamda(cutneigh, c);
    const int k[3] = {(int)std::ceil(c[0]), (int)std::ceil(c[1]), (int)std::ceil(c[2])};
    x.resize(nlocal);
    type.resize(nlocal);
    owner.clear();
    image.clear();
    for (int nz = -k[2]; nz <= k[2]; ++nz)
      for (int ny = -k[1]; ny <= k[1]; ++ny)
        for (int nx = -k[0]; nx <= k[0]; ++nx) {
          if (nx == 0 && ny == 0 && nz == 0) continue;
          const int n[3] = {nx, ny, nz};
          for (int i = 0; i < nlocal; ++i) {
            double l[3];
            dom.x2lamda(x[i].data(), l);
            bool in = true;
            for (int d = 0; d < 3; ++d) {
              const double s = l[d] + n[d];
              in = in && s >= -c[d] && s < 1.0 + c[d];
            }
            if (!in) continue;
            double sh[3];
            dom.shift(n, sh);
            x.push_back({x[i][0] + sh[0], x[i][1] + sh[1], x[i][2] + sh[2]});
            type.push_back(type[i]);
            owner.push_back(i);
            image.push_back({nx, ny, nz});
          }
        }
    nghost = (int)owner.size();
    f.assign(nlocal + nghost, {0, 0, 0});
  }

  void forward_comm() {
    for (int g = 0; g < nghost; ++g) {
      double sh[3];
      dom.shift(image[g].data(), sh);
      const auto& o = x[owner[g]];
      x[nlocal + g] = {o[0] + sh[0], o[1] + sh[1], o[2] + sh[2]};
    }
  }

  void build() {
    const int nall = nlocal + nghost;
    const double c2 = cutneigh * cutneigh;
    ilist.resize(nlocal);
    numneigh.resize(nlocal);
    rows.assign(nlocal, {});
    for (int ii = 0; ii < nlocal; ++ii) {
      // LAMMPS orders ilist by its own bins; any order is allowed
      const int i = shuffle ? (nlocal - 1 - ii) : ii;
      ilist[ii] = i;
      auto& row = rows[ii];
      for (int j = 0; j < nall; ++j) {
        if (j == i) continue;
        double r2 = 0;
        for (int d = 0; d < 3; ++d) r2 += (x[j][d] - x[i][d]) * (x[j][d] - x[i][d]);
        if (r2 < c2) row.push_back(j);
      }
      if (shuffle)
        for (int k = (int)row.size() - 1; k > 0; --k) std::swap(row[k], row[next_rand() % (k + 1)]);
      if (special)
        for (size_t k = 0; k < row.size(); k += 3) row[k] |= 1 << 30;
      numneigh[ii] = (int)row.size();
    }
    first.resize(nlocal);
    for (int ii = 0; ii < nlocal; ++ii) first[ii] = rows[ii].data();
    xhold.assign(x.begin(), x.begin() + nlocal);
    ago = 0;
    ++builds;
  }

  bool need_build() const {
    const double lim = 0.25 * skin * skin;
    for (int i = 0; i < nlocal; ++i) {
      double r2 = 0;
      for (int d = 0; d < 3; ++d) r2 += (x[i][d] - xhold[i][d]) * (x[i][d] - xhold[i][d]);
      if (r2 > lim) return true;
    }
    return false;
  }

  dpmi::lmp::System system() {
    const int nall = nlocal + nghost;
    xp.resize(nall);
    fp.resize(nall);
    for (int i = 0; i < nall; ++i) {
      xp[i] = x[i].data();
      fp[i] = f[i].data();
    }
    dpmi::lmp::System s;
    s.x = xp.data();
    s.f = fp.data();
    s.type = type.data();
    s.nlocal = nlocal;
    s.nghost = nghost;
    s.ntypes = ntypes;
    std::copy(dom.boxlo, dom.boxlo + 3, s.boxlo);
    std::copy(dom.h, dom.h + 6, s.h);
    s.inum = nlocal;
    s.ilist = ilist.data();
    s.numneigh = numneigh.data();
    s.firstneigh = first.data();
    s.ago = ago;
    return s;
  }

  void reverse_comm() {
    for (int g = 0; g < nghost; ++g)
      for (int d = 0; d < 3; ++d) f[owner[g]][d] += f[nlocal + g][d];
  }

  double kinetic() const {
    double ke = 0;
    for (int i = 0; i < nlocal; ++i)
      ke += 0.5 * mass[type[i] - 1] * (v[i][0] * v[i][0] + v[i][1] * v[i][1] + v[i][2] * v[i][2]);
    return ke * kMvv2e;
  }
};

int run_errors(const std::string& model) {
  std::vector<int> raised;
  auto expect = [&](auto fn) {
    try {
      fn();
      raised.push_back(0);
    } catch (const deepmd::deepmd_exception&) {
      raised.push_back(1);
    }
  };
  dpmi::lmp::PairDeepMD pair;
  expect([&] { pair.settings({}); });
  expect([&] { pair.coeff({"*", "*"}, 2); });  // before pair_style
  pair.settings({model});
  expect([&] { pair.coeff({"1", "2"}, 2); });
  expect([&] { pair.coeff({"*", "*", "O", "Xx"}, 2); });
  expect([&] { pair.settings({model, "precision"}); });
  pair.settings({model});
  pair.coeff({"*", "*", "O", "H"}, 2);
  double xs[2][3] = {{0, 0, 0}, {1, 0, 0}}, fs[2][3] = {};
  double* xptr[2] = {xs[0], xs[1]};
  double* fptr[2] = {fs[0], fs[1]};
  int types[2] = {1, 2}, il[2] = {0, 1}, nn[2] = {1, 1}, r0[1] = {1}, r1[1] = {0};
  int* fn[2] = {r0, r1};
  dpmi::lmp::System s;
  s.x = xptr;
  s.f = fptr;
  s.type = types;
  s.nlocal = 2;
  s.ntypes = 2;
  s.h[0] = s.h[1] = s.h[2] = 20.0;
  s.inum = 2;
  s.ilist = il;
  s.numneigh = nn;
  s.firstneigh = fn;
  expect([&] { pair.compute(s, 1, 1, false, true, false); });  // 6-component per-atom virial
  types[1] = 3;
  expect([&] { pair.compute(s, 1, 1); });  // LAMMPS type beyond pair_coeff
  types[1] = 2;
  expect([&] { pair.compute(s, 1, 1); });  // a valid call raises nothing
  wr("errors.i32", raised);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc < 8) {
      std::cerr << "usage: dpmi_md <dir> <model> <precision> <steps> <dt> <skin> <dump_every> [options]\n";
      return 2;
    }
    g_dir = argv[1];
    const std::string model = argv[2], precision = argv[3];
    const int nsteps = std::stoi(argv[4]);
    const double dt = std::stod(argv[5]);
    const double skin = std::stod(argv[6]);
    const int dump_every = std::stoi(argv[7]);
    std::vector<std::string> coeff = {"*", "*"};
    bool shuffle = false, special = false, atom = false;
    for (int k = 8; k < argc; ++k) {
      const std::string a = argv[k];
      if (a == "errors") return run_errors(model);
      if (a == "shuffle") shuffle = true;
      if (a == "special") special = true;
      if (a == "atom") atom = true;
      if (a == "names")
        while (k + 1 < argc && std::string(argv[k + 1]) != "shuffle" && std::string(argv[k + 1]) != "special" &&
               std::string(argv[k + 1]) != "atom")
          coeff.push_back(argv[++k]);
    }

    Md md;
    md.dom.set(rd<double>("box.f64"), rd<double>("boxlo.f64"));
    const std::vector<double> c = rd<double>("coord.f64"), vel = rd<double>("vel.f64");
    md.type = rd<int>("type.i32");
    md.mass = rd<double>("mass.f64");
    md.nlocal = (int)md.type.size();
    md.ntypes = (int)md.mass.size();
    for (int i = 0; i < md.nlocal; ++i) {
      md.x.push_back({c[3 * i], c[3 * i + 1], c[3 * i + 2]});
      md.v.push_back({vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]});
    }
    md.skin = skin;
    md.shuffle = shuffle;
    md.special = special;

    dpmi::lmp::PairDeepMD pair;
    pair.settings({model, "precision", precision});
    pair.coeff(coeff, md.ntypes);
    md.cutneigh = pair.init_one(1, 1) + skin;

    std::vector<double> thermo;
    auto force_step = [&](int step) {
      for (auto& fi : md.f) fi = {0, 0, 0};
      const bool dump = dump_every > 0 && step % dump_every == 0;
      const bool at = atom && dump;
      dpmi::lmp::System s = md.system();
      pair.compute(s, 1, 1, at, false, at);
      md.reverse_comm();
      if (dump) {
        const std::string p = "dump" + std::to_string(step) + "_";
        std::vector<double> xs, fs;
        for (int i = 0; i < md.nlocal; ++i)
          for (int d = 0; d < 3; ++d) {
            xs.push_back(md.x[i][d]);
            fs.push_back(md.f[i][d]);
          }
        wr(p + "x.f64", xs);
        wr(p + "f.f64", fs);
        if (at) {
          std::vector<double> ea(pair.eatom.begin(), pair.eatom.begin() + md.nlocal);
          std::vector<std::array<double, 9>> cv(pair.cvatom.begin(), pair.cvatom.begin() + md.nlocal);
          for (int g = 0; g < md.nghost; ++g)
            for (int k = 0; k < 9; ++k) cv[md.owner[g]][k] += pair.cvatom[md.nlocal + g][k];
          std::vector<double> cvf;
          for (auto& r : cv) cvf.insert(cvf.end(), r.begin(), r.end());
          wr(p + "eatom.f64", ea);
          wr(p + "cvatom.f64", cvf);
        }
      }
    };
    auto record = [&](int step) {
      const double ke = md.kinetic();
      const double row[12] = {(double)step, pair.eng_vdwl, ke, pair.eng_vdwl + ke, (double)md.ago, (double)md.builds,
                              pair.virial[0], pair.virial[1], pair.virial[2], pair.virial[3], pair.virial[4],
                              pair.virial[5]};
      thermo.insert(thermo.end(), row, row + 12);
    };

    // Verlet::setup
    md.pbc();
    md.borders();
    md.build();
    force_step(0);
    record(0);
    const double dtf = 0.5 * dt * kFtm2v;
    for (int step = 1; step <= nsteps; ++step) {
      for (int i = 0; i < md.nlocal; ++i) {
        const double m = md.mass[md.type[i] - 1];
        for (int d = 0; d < 3; ++d) {
          md.v[i][d] += dtf * md.f[i][d] / m;
          md.x[i][d] += dt * md.v[i][d];
        }
      }
      ++md.ago;
      if (md.need_build()) {
        md.pbc();
        md.borders();
        md.build();
      } else {
        md.forward_comm();
      }
      force_step(step);
      for (int i = 0; i < md.nlocal; ++i) {
        const double m = md.mass[md.type[i] - 1];
        for (int d = 0; d < 3; ++d) md.v[i][d] += dtf * md.f[i][d] / m;
      }
      record(step);
      if (step % 100 == 0) std::printf("step %d pe %.10f etot %.10f builds %d\n", step, pair.eng_vdwl,
                                       thermo[thermo.size() - 9], md.builds);
      std::fflush(stdout);
    }
    wr("thermo.f64", thermo);
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "dpmi_md: " << e.what() << "\n";
    return 1;
  }
}
