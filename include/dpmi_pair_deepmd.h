// dpmi::lmp::PairDeepMD -- the LAMMPS pair style "deepmd" over the MI355X
// se_e2_a path, with the per-step contract of the reference's PairDeepMD
// (source/lmp/pair_deepmd.cpp:253-632, coeff :1013-1100, unit factors
// pair_base.cpp:391-399, full neighbour request pair_base.cpp:544-551).
//
// LAMMPS is not linked: the pair style reads the arrays LAMMPS hands a pair
// style (Atom, Domain, Force, NeighList, Neighbor::ago) from a plain struct,
// so the same code runs under LAMMPS (a PairStyle subclass that fills System
// from its members and forwards settings/coeff/init_one/compute) and under a
// test driver.  Per compute():
//   * LAMMPS types 1..n map to model types by pair_coeff names (NULL = -1,
//     an absent atom);
//   * coordinates are shifted by boxlo and converted from LAMMPS distance
//     units; the box is rebuilt from domain->h;
//   * the full neighbour list (cutoff + skin, ghosts included, any order) is
//     passed with NEIGHMASK applied; ago == 0 uploads it, ago > 0 reuses it;
//   * forces are added onto all nlocal + nghost atoms: LAMMPS's reverse
//     communication (newton on) folds the ghost terms onto their owners;
//   * energy, the 6-component global virial, per-atom energy (local atoms)
//     and the 9-component centroid per-atom virial (all atoms) are
//     accumulated as the reference does; a 6-component per-atom virial
//     raises, as in the reference.
// Errors are deepmd::deepmd_exception (LAMMPS would pass e.what() to
// error->one).
#pragma once

#include <array>
#include <string>
#include <vector>

#include "dpmi_deeppot.h"

namespace dpmi {
namespace lmp {

// LAMMPS keeps special-bond bits above this mask in neighbour indices.
constexpr int NEIGHMASK = 0x1FFFFFFF;

// The LAMMPS state one pair-style compute() reads and writes.
struct System {
  double** x = nullptr;      // atom->x [nlocal + nghost][3]
  double** f = nullptr;      // atom->f [nlocal + nghost][3], accumulated into
  const int* type = nullptr;  // atom->type, 1..ntypes
  int nlocal = 0, nghost = 0, ntypes = 0;
  double boxlo[3] = {0, 0, 0};
  double h[6] = {0, 0, 0, 0, 0, 0};  // domain->h: xx yy zz yz xz xy
  // the pair style's full neighbour list (list->inum, ilist, numneigh,
  // firstneigh) and neighbor->ago
  int inum = 0;
  int* ilist = nullptr;
  int* numneigh = nullptr;
  int** firstneigh = nullptr;
  int ago = 0;
  // force->boltz and force->angstrom (metal units by default)
  double boltz = 8.617343e-5;
  double angstrom = 1.0;
  // fix adapt's scale[1][1]
  double scale = 1.0;
};

class PairDeepMD {
 public:
  PairDeepMD() = default;
  PairDeepMD(const PairDeepMD&) = delete;
  PairDeepMD& operator=(const PairDeepMD&) = delete;

  // pair_style deepmd <model file> [precision fp64|mixed] [gpu <rank>]
  void settings(const std::vector<std::string>& args);
  // pair_coeff * * [type names...]; ntypes = atom->ntypes
  void coeff(const std::vector<std::string>& args, int ntypes);
  // the model cutoff in LAMMPS distance units (init_one for every type pair)
  double init_one(int i, int j) const;

  // ev_init semantics: the accumulators below are reset when their flag is
  // set, then this step's terms are added
  void compute(System& s, int eflag, int vflag, bool eflag_atom = false, bool vflag_atom = false,
               bool cvflag_atom = false);

  double eng_vdwl = 0.0;
  double virial[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
  std::vector<double> eatom;                 // [nlocal + nghost]
  std::vector<std::array<double, 9>> cvatom;  // [nlocal + nghost], xx yy zz xy xz yz yx zx zy

  const std::vector<int>& type_idx_map() const { return type_idx_map_; }

 private:
  DeepPot deep_pot_;
  bool ready_ = false;
  std::vector<int> type_idx_map_;
  double cutoff_ = 0.0;
};

}  // namespace lmp
}  // namespace dpmi
