// dpmi::DeepPot -- the DeepPot-style C++ model backend of the MI355X se_e2_a
// path, with the evaluation methods of the reference's DeepPot
// (source/api_cc/include/DeepPot.h:456-700, implemented for PyTorch models in
// source/api_cc/src/DeepPotPT.cc:171-500):
//   compute(ener, force, virial, coord, atype, box)                      no neighbour list
//   compute(ener, force, virial, atom_energy, atom_virial, coord, atype, box)
//   compute(ener, force, virial, coord, atype, box, nghost, inlist, ago)  caller's ghosts and list
//   compute(ener, force, virial, atom_energy, atom_virial, coord, atype, box, nghost, inlist, ago)
// Host vectors in and out, as in the reference; the work runs on one GPU
// (gpu_rank) through the dpmi_sea_* engine of include/dpmi.h:
// * without a list, a periodic box (box.size() == 9) goes through the
//   engine's own cell list and ghost images; an empty box means no periodic
//   images, and the candidate rows are all pairs within rcut;
// * with a list (LAMMPS, pair_deepmd.cpp:253-391) coord/atype hold nall =
//   nloc + nghost atoms, inlist holds the rows of the nloc local atoms (any
//   order, rcut + skin allowed); with ago == 0 the rows are uploaded, else the
//   last upload is reused.  force and atom_virial are returned for all nall
//   atoms (ghost terms not folded: the MD code sums them onto their owners);
//   atom_energy has nall entries, zero for ghosts.
// Atoms of negative type are absent: no neighbours, no energy, zero force.
// fparam / aparam are not used by se_e2_a energy models here: a non-empty one
// raises.  Errors are deepmd::deepmd_exception (include/deepmd_lib.h).
#pragma once

#include <string>
#include <vector>

#include "deepmd_lib.h"

namespace dpmi {

class DeepPot {
 public:
  DeepPot();
  // model: a dpmi model file (dpmi_sea_model_load); precision "fp64" (the
  // plain path, parity with the reference) or "mixed"
  explicit DeepPot(const std::string& model, const int& gpu_rank = 0, const std::string& precision = "fp64");
  ~DeepPot();
  DeepPot(const DeepPot&) = delete;
  DeepPot& operator=(const DeepPot&) = delete;
  void init(const std::string& model, const int& gpu_rank = 0, const std::string& precision = "fp64");

  template <typename VALUETYPE>
  void compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
               const std::vector<VALUETYPE>& coord, const std::vector<int>& atype, const std::vector<VALUETYPE>& box,
               const std::vector<VALUETYPE>& fparam = std::vector<VALUETYPE>(),
               const std::vector<VALUETYPE>& aparam = std::vector<VALUETYPE>());
  template <typename VALUETYPE>
  void compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
               std::vector<VALUETYPE>& atom_energy, std::vector<VALUETYPE>& atom_virial,
               const std::vector<VALUETYPE>& coord, const std::vector<int>& atype, const std::vector<VALUETYPE>& box,
               const std::vector<VALUETYPE>& fparam = std::vector<VALUETYPE>(),
               const std::vector<VALUETYPE>& aparam = std::vector<VALUETYPE>());
  template <typename VALUETYPE>
  void compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
               const std::vector<VALUETYPE>& coord, const std::vector<int>& atype, const std::vector<VALUETYPE>& box,
               const int nghost, const deepmd::InputNlist& inlist, const int& ago,
               const std::vector<VALUETYPE>& fparam = std::vector<VALUETYPE>(),
               const std::vector<VALUETYPE>& aparam = std::vector<VALUETYPE>());
  template <typename VALUETYPE>
  void compute(double& ener, std::vector<VALUETYPE>& force, std::vector<VALUETYPE>& virial,
               std::vector<VALUETYPE>& atom_energy, std::vector<VALUETYPE>& atom_virial,
               const std::vector<VALUETYPE>& coord, const std::vector<int>& atype, const std::vector<VALUETYPE>& box,
               const int nghost, const deepmd::InputNlist& inlist, const int& ago,
               const std::vector<VALUETYPE>& fparam = std::vector<VALUETYPE>(),
               const std::vector<VALUETYPE>& aparam = std::vector<VALUETYPE>());

  double cutoff() const;
  int numb_types() const;
  int dim_fparam() const { return 0; }
  int dim_aparam() const { return 0; }
  void get_type_map(std::string& type_map);

 private:
  struct Impl;
  Impl* impl_ = nullptr;
};

}  // namespace dpmi
