// The LAMMPS pair style "deepmd" over dpmi::DeepPot (declared in
// include/dpmi_pair_deepmd.h).  Reference: source/lmp/pair_deepmd.cpp.
#include "../../include/dpmi_pair_deepmd.h"

#include <sstream>
#include <stdexcept>

namespace dpmi {
namespace lmp {

namespace {
std::vector<std::string> split_words(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> out;
  for (std::string w; is >> w;) out.push_back(w);
  return out;
}
}  // namespace

void PairDeepMD::settings(const std::vector<std::string>& args) {
  if (args.empty()) throw deepmd::deepmd_exception("pair_style deepmd needs a model file");
  std::string precision = "fp64";
  int gpu = 0;
  for (size_t k = 1; k < args.size(); k += 2) {
    if (k + 1 >= args.size()) throw deepmd::deepmd_exception("pair_style deepmd: keyword " + args[k] + " needs a value");
    if (args[k] == "precision")
      precision = args[k + 1];
    else if (args[k] == "gpu")
      gpu = std::stoi(args[k + 1]);
    else
      throw deepmd::deepmd_exception("pair_style deepmd: unknown keyword " + args[k]);
  }
  deep_pot_.init(args[0], gpu, precision);
  cutoff_ = deep_pot_.cutoff();
  ready_ = true;
  type_idx_map_.clear();
}

// pair_deepmd.cpp:1013-1100: "* *" only; names map LAMMPS types onto the
// model's type_map, NULL marks an absent atom; without names the mapping is
// positional.  Unnamed trailing LAMMPS types map positionally too.
void PairDeepMD::coeff(const std::vector<std::string>& args, int ntypes) {
  if (!ready_) throw deepmd::deepmd_exception("pair_coeff before pair_style deepmd");
  if (ntypes <= 0) throw deepmd::deepmd_exception("pair_coeff: the system has no atom types");
  if (args.size() >= 2 && (args[0] != "*" || args[1] != "*"))
    throw deepmd::deepmd_exception(
        "deepmd requires that the scale should be set to all atom types, i.e. pair_coeff * *.");
  type_idx_map_.clear();
  if (args.size() <= 2) {
    for (int i = 0; i < ntypes; ++i) type_idx_map_.push_back(i);
    return;
  }
  std::string tm;
  deep_pot_.get_type_map(tm);
  const std::vector<std::string> type_map = split_words(tm);
  for (size_t k = 2; k < args.size(); ++k) {
    int idx = -2;
    for (size_t t = 0; t < type_map.size(); ++t)
      if (type_map[t] == args[k]) {
        idx = (int)t;
        break;
      }
    if (idx == -2 && args[k] == "NULL") idx = -1;
    if (idx == -2) throw deepmd::deepmd_exception("pair_coeff: element " + args[k] + " not found in the model");
    type_idx_map_.push_back(idx);
  }
  for (int i = (int)type_idx_map_.size(); i < ntypes; ++i) type_idx_map_.push_back(i);
  type_idx_map_.resize(ntypes);
}

double PairDeepMD::init_one(int, int) const {
  if (!ready_) throw deepmd::deepmd_exception("pair style deepmd is not set up");
  return cutoff_;
}

// pair_deepmd.cpp:253-391 (single model, ghosts on) and :612-632
void PairDeepMD::compute(System& s, int eflag, int vflag, bool eflag_atom, bool vflag_atom, bool cvflag_atom) {
  if (!ready_) throw deepmd::deepmd_exception("pair style deepmd is not set up");
  if (type_idx_map_.empty()) throw deepmd::deepmd_exception("pair_coeff is missing for pair style deepmd");
  if (vflag_atom)
    throw deepmd::deepmd_exception(
        "6-element atomic virial is not supported. Use compute centroid/stress/atom command for 9-element atomic "
        "virial.");
  const int nlocal = s.nlocal, nall = s.nlocal + s.nghost;
  const double ener_cvt = s.boltz / 8.617343e-5, dist_cvt = s.angstrom, force_cvt = ener_cvt / dist_cvt;
  // ev_init
  if (eflag) eng_vdwl = 0.0;
  if (vflag)
    for (double& v : virial) v = 0.0;
  if (eflag_atom) eatom.assign(nall, 0.0);
  if (cvflag_atom) cvatom.assign(nall, std::array<double, 9>{});

  std::vector<int> dtype(nall);
  for (int i = 0; i < nall; ++i) {
    const int t = s.type[i];
    if (t < 1 || t > (int)type_idx_map_.size()) throw deepmd::deepmd_exception("atom type out of the pair_coeff range");
    dtype[i] = type_idx_map_[t - 1];
  }
  std::vector<double> dbox(9, 0.0), dcoord(3 * (size_t)nall);
  dbox[0] = s.h[0] / dist_cvt;
  dbox[4] = s.h[1] / dist_cvt;
  dbox[8] = s.h[2] / dist_cvt;
  dbox[7] = s.h[3] / dist_cvt;
  dbox[6] = s.h[4] / dist_cvt;
  dbox[3] = s.h[5] / dist_cvt;
  for (int i = 0; i < nall; ++i)
    for (int d = 0; d < 3; ++d) dcoord[3 * (size_t)i + d] = (s.x[i][d] - s.boxlo[d]) / dist_cvt;

  deepmd::InputNlist list(s.inum, s.ilist, s.numneigh, s.firstneigh);
  list.set_mask(NEIGHMASK);
  double dener = 0.0;
  std::vector<double> dforce, dvirial;
  if (!(eflag_atom || cvflag_atom)) {
    deep_pot_.compute(dener, dforce, dvirial, dcoord, dtype, dbox, s.nghost, list, s.ago);
  } else {
    std::vector<double> deatom, dvatom;
    deep_pot_.compute(dener, dforce, dvirial, deatom, dvatom, dcoord, dtype, dbox, s.nghost, list, s.ago);
    if (eflag_atom)
      for (int i = 0; i < nlocal; ++i) eatom[i] += s.scale * deatom[i] * ener_cvt;
    if (cvflag_atom) {
      // DeePMD's row-major 3x3 onto LAMMPS's centroid order
      static const int order[9] = {0, 4, 8, 3, 6, 7, 1, 2, 5};
      for (int i = 0; i < nall; ++i)
        for (int k = 0; k < 9; ++k) cvatom[i][k] += s.scale * dvatom[9 * (size_t)i + order[k]] * ener_cvt;
    }
  }
  for (int i = 0; i < nall; ++i)
    for (int d = 0; d < 3; ++d) s.f[i][d] += s.scale * dforce[3 * (size_t)i + d] * force_cvt;
  if (eflag) eng_vdwl += s.scale * dener * ener_cvt;
  if (vflag) {
    virial[0] += dvirial[0] * s.scale * ener_cvt;
    virial[1] += dvirial[4] * s.scale * ener_cvt;
    virial[2] += dvirial[8] * s.scale * ener_cvt;
    virial[3] += dvirial[3] * s.scale * ener_cvt;
    virial[4] += dvirial[6] * s.scale * ener_cvt;
    virial[5] += dvirial[7] * s.scale * ener_cvt;
  }
}

}  // namespace lmp
}  // namespace dpmi
