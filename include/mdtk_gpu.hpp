// mdtk_gpu.hpp - C++ host adapter over the C ABI of mdg.h, shaped like the
// reference's potential interface so that it drops into an mdtrajsim-style
// driver.
//
// PotentialChain plays the role of one FProxy holding the chain
// FBZL -> REBO -> AIREBO -> ETors (applications/common.h:46-68) with the
// FGeneral contract (mdtk/potentials/FGeneral.hpp:54-114):
//   * operator()(atoms) returns the energy (erg) and ADDS dE/dx (erg/cm) to
//     every Atom::grad, like FGeneral::operator();
//   * getRcutoff() equals the largest member cutoff, as FProxy::getRcutoff
//     (FProxy.cxx:35-43) reports it, so checkMIC behaves identically;
//   * incDisplacement(atom, inc) feeds the Verlet-skin test exactly as
//     FGeneral::incDisplacement does (FGeneral.hpp:87-91);
//   * hasNeighbors(atom) answers FProxy::hasNeighbors (FProxy.hpp:99-104).
// Errors are returned by the C ABI as status codes and rethrown here as
// mdtk_gpu::Error, so the reference's crash handling (SimLoop.cxx:176-205)
// keeps working.
//
// The template parameter is the caller's atom container (mdtk::AtomsArray):
// it needs size(), operator[] and, per atom, ID, coords.{x,y,z},
// PBC.{x,y,z}, PBC_count.{x,y,z}, grad.{x,y,z}, globalIndex, V.{x,y,z} and
// isFixed() -- the fields of mdtk::Atom (mdtk/Atom.hpp:32-96).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdg.h"

namespace mdtk_gpu {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class AtomsArray>
class PotentialChain {
 public:
  explicit PotentialChain(int pot_mask = MDG_POT_FBZL | MDG_POT_REBO | MDG_POT_AIREBO | MDG_POT_ETORS,
                          bool etors_compat = true, int brenner_paramset = 0, double fbzl_r0 = 5.0e-8,
                          double fbzl_r1 = 5.5e-8) {
    mdg_config cfg{};
    cfg.pot_mask = pot_mask;
    cfg.brenner_paramset = brenner_paramset;
    cfg.etors_compat = etors_compat ? 1 : 0;
    cfg.precision = MDG_PREC_FP64;
    cfg.fbzl_r0 = fbzl_r0;
    cfg.fbzl_r1 = fbzl_r1;
    if (mdg_create(&cfg, &ctx_) != 0) throw Error(std::string("mdg_create: ") + mdg_last_error(nullptr));
    rc_ = 0.0;
    if (pot_mask & (MDG_POT_REBO | MDG_POT_BRENNER | MDG_POT_ETORS)) rc_ = 2.0e-8;  // REBO.hpp:130
    if (pot_mask & MDG_POT_AIREBO) rc_ = 5.5e-8;                                    // AIREBO_LJ.hpp:81
    if ((pot_mask & MDG_POT_FBZL) && fbzl_r1 > rc_) rc_ = fbzl_r1;                  // FPairwise.hpp:62
  }
  ~PotentialChain() { mdg_destroy(ctx_); }
  PotentialChain(const PotentialChain&) = delete;
  PotentialChain& operator=(const PotentialChain&) = delete;

  double getRcutoff() const { return rc_; }

  // FGeneral::incDisplacement: accumulated until the next force call.
  void incDisplacement(std::size_t globalIndex, double dx, double dy, double dz) {
    if (dr_.size() <= 3 * globalIndex + 2) return;
    dr_[3 * globalIndex] += dx;
    dr_[3 * globalIndex + 1] += dy;
    dr_[3 * globalIndex + 2] += dz;
    have_dr_ = true;
  }

  // FGeneral::operator(): energy in erg; adds dE/dx into each atom's grad.
  double operator()(AtomsArray& atoms) {
    const std::size_t n = atoms.size();
    pack(atoms);
    if (!ready_) {
      std::vector<int32_t> elem(n);
      std::vector<double> vel(3 * n);
      std::vector<uint8_t> fixed(n);
      for (std::size_t i = 0; i < n; i++) {
        elem[i] = static_cast<int32_t>(atoms[i].ID);
        vel[3 * i] = atoms[i].V.x;
        vel[3 * i + 1] = atoms[i].V.y;
        vel[3 * i + 2] = atoms[i].V.z;
        fixed[i] = atoms[i].isFixed() ? 1 : 0;
      }
      // the box: per axis the periodic value the atoms carry (the array's
      // PBC); atoms open on a periodic axis keep their own PBC vectors
      // through mdg_set_atom_pbc (Atom.hpp:115-140)
      const double no_pbc = 10000.0 * 1e-8;  // NO_PBC, mdtk/consts.hpp:96
      double pbc[3] = {no_pbc, no_pbc, no_pbc};
      std::vector<double> own(3 * n);
      for (std::size_t i = 0; i < n; i++) {
        own[3 * i] = atoms[i].PBC.x;
        own[3 * i + 1] = atoms[i].PBC.y;
        own[3 * i + 2] = atoms[i].PBC.z;
        for (int d = 0; d < 3; d++)
          if (own[3 * i + d] != no_pbc) pbc[d] = own[3 * i + d];
      }
      bool per_atom = false;
      for (std::size_t q = 0; q < 3 * n; q++) per_atom = per_atom || own[q] != pbc[q % 3];
      check(mdg_set_system(ctx_, static_cast<int>(n), elem.data(), xyz_.data(), vel.data(), pbc, fixed.data()));
      if (per_atom) check(mdg_set_atom_pbc(ctx_, own.data(), xyz_.data(), pc_.data()));
      else check(mdg_set_positions(ctx_, xyz_.data(), pc_.data(), nullptr));
      dr_.assign(3 * n, 0.0);
      ready_ = true;
    } else {
      check(mdg_set_positions(ctx_, xyz_.data(), pc_.data(), have_dr_ ? dr_.data() : nullptr));
      std::fill(dr_.begin(), dr_.end(), 0.0);
      have_dr_ = false;
    }
    e_.assign(static_cast<std::size_t>(mdg_num_terms(ctx_)), 0.0);
    grad_.assign(3 * n, 0.0);
    check(mdg_energy_grad(ctx_, e_.data(), grad_.data()));
    double E = 0.0;
    for (double v : e_) E += v;
    for (std::size_t i = 0; i < n; i++) {
      atoms[i].grad.x += grad_[3 * i];
      atoms[i].grad.y += grad_[3 * i + 1];
      atoms[i].grad.z += grad_[3 * i + 2];
    }
    counts_.assign(n, 0);
    check(mdg_neighbor_counts(ctx_, counts_.data()));
    return E;
  }

  // Energies of the chain members of the last call (the E<i> lines of
  // FProxy.cxx:57), in chain order.
  const std::vector<double>& lastEnergies() const { return e_; }

  bool hasNeighbors(std::size_t globalIndex) const {
    return globalIndex < counts_.size() && counts_[globalIndex] > 0;
  }

  mdg_ctx* context() { return ctx_; }

 private:
  void check(int rc) {
    if (rc != 0) throw Error(mdg_last_error(ctx_));
  }
  void pack(AtomsArray& atoms) {
    const std::size_t n = atoms.size();
    xyz_.resize(3 * n);
    pc_.resize(3 * n);
    for (std::size_t i = 0; i < n; i++) {
      xyz_[3 * i] = atoms[i].coords.x;
      xyz_[3 * i + 1] = atoms[i].coords.y;
      xyz_[3 * i + 2] = atoms[i].coords.z;
      pc_[3 * i] = atoms[i].PBC_count.x;
      pc_[3 * i + 1] = atoms[i].PBC_count.y;
      pc_[3 * i + 2] = atoms[i].PBC_count.z;
    }
  }

  mdg_ctx* ctx_ = nullptr;
  double rc_ = 0.0;
  bool ready_ = false;
  bool have_dr_ = false;
  std::vector<double> xyz_, dr_, grad_, e_;
  std::vector<int32_t> pc_, counts_;
};

}  // namespace mdtk_gpu
