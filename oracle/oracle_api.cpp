// C entry points of the oracle, for tests and the CPU-baseline leg of the
// benchmark.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Units are the reference's CGS units (cm, cm/s, erg).  A non-periodic axis is
// given as NO_PBC = 1e4 Angstrom = 1e-4 cm, as in mdtk/consts.hpp:96.
#include <cstring>
#include <string>

#include "oracle_core.hpp"

using namespace orc;

namespace {
thread_local std::string g_err;

enum : int { POT_FBZL = 1, POT_BRENNER = 2, POT_REBO = 4, POT_AIREBO = 8, POT_ETORS = 16, POT_FLJ = 32, POT_ACKLAND = 64,
            POT_FBM = 128, POT_TB = 256 };

struct Handle {
  SimLoop sl;
  std::vector<int> kinds;  // kind of each chain entry, in chain order
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_alter_exact(int on) { g_alter_exact = on != 0; }

// Chain order follows applications/common.h:46-68: FBZL, (FLJ: inert for H/C/Ar),
// REBO, AIREBO(LJ), ETors, Ackland.  Brenner replaces REBO when selected.
void* orc_create(int n, const int* elem, const double* xyz, const double* vel, const double* pbc,
                 const unsigned char* fixed, int pot_mask, int brenner_set, int etors_compat,
                 double fbzl_r0, double fbzl_r1) {
  Handle* h = nullptr;
  int rc = guarded([&] {
    h = new Handle();
    auto& at = h->sl.atoms;
    at.resize(n);
    for (int i = 0; i < n; i++) {
      at[i].ID = elem[i];
      at[i].coords = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
      if (vel) at[i].V = V3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
      at[i].PBC = V3(pbc[0], pbc[1], pbc[2]);
      at[i].fixed = fixed ? (fixed[i] != 0) : false;
      if (at[i].fixed) at[i].V = V3();
      at[i].globalIndex = i;
      at[i].setAttributesByElementID();
    }
    h->sl.arrayPBC = V3(pbc[0], pbc[1], pbc[2]);
    auto& pots = h->sl.fpot.pots;
    if (pot_mask & POT_FBZL) {
      pots.push_back(new FBZL(fbzl_r0, fbzl_r1));
      h->kinds.push_back(POT_FBZL);
    }
    if (pot_mask & POT_FLJ) {  // common.h:49: FLJ(Rcutoff(5.0 A, 5.5 A)) after FBZL
      pots.push_back(new FLJ(5.0 * Ao, 5.5 * Ao));
      h->kinds.push_back(POT_FLJ);
    }
    REBO* rebo = nullptr;
    if (pot_mask & POT_BRENNER) {
      pots.push_back(new Brenner(brenner_set));
      h->kinds.push_back(POT_BRENNER);
    }
    if (pot_mask & (POT_REBO | POT_AIREBO)) {
      rebo = new REBO();
      pots.push_back(rebo);
      h->kinds.push_back(POT_REBO);
    }
    if (pot_mask & POT_AIREBO) {
      pots.push_back(new AIREBO_LJ(*rebo));
      h->kinds.push_back(POT_AIREBO);
    }
    if (pot_mask & POT_ETORS) {
      auto* t = new ETors();
      t->compat = etors_compat != 0;
      pots.push_back(t);
      h->kinds.push_back(POT_ETORS);
    }
    if (pot_mask & POT_ACKLAND) {  // common.h:67: Ackland last
      pots.push_back(new Ackland());
      h->kinds.push_back(POT_ACKLAND);
    }
    // FBM and TightBinding are not registered by common.h (:69-72); selected,
    // they follow the chain in this order.
    if (pot_mask & POT_FBM) {  // FPairwise's default Rcutoff (FPairwise.hpp:34): 3.1-3.5 A
      pots.push_back(new FBM(3.1 * Ao, 3.5 * Ao));
      h->kinds.push_back(POT_FBM);
    }
    if (pot_mask & POT_TB) {
      pots.push_back(new TightBinding());
      h->kinds.push_back(POT_TB);
    }
    if (pots.empty()) throw OracleError("no potentials selected");
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return h;
}

void orc_destroy(void* p) { delete static_cast<Handle*>(p); }

int orc_num_terms(void* p) { return (int)static_cast<Handle*>(p)->kinds.size(); }
int orc_term_kind(void* p, int i) { return static_cast<Handle*>(p)->kinds[i]; }

// The virial of one force call at the current positions (six components,
// erg), from the pushes of every primitive (see g_vir in oracle_core.hpp).
int orc_virial(void* p, double* out) {
  Handle* h = static_cast<Handle*>(p);
  return guarded([&] {
    auto& sl = h->sl;
    if (!sl.prepared) sl.prepare();
    sl.fpot.NL_UpdateIfNeeded(sl.atoms);
    for (double& v : g_vir) v = 0.0;
    g_vir_on = true;
    try {
      sl.fpot.compute(sl.atoms);
    } catch (...) {
      g_vir_on = false;
      throw;
    }
    g_vir_on = false;
    for (int c = 0; c < 6; c++) out[c] = g_vir[c];
  });
}

// Energy (per chain entry, erg) and dE/dx (erg/cm) at the current positions,
// exactly as one FProxy call made from SimLoop::energy().
int orc_energy_grad(void* p, double* e_terms, double* grad) {
  Handle* h = static_cast<Handle*>(p);
  return guarded([&] {
    auto& sl = h->sl;
    if (!sl.prepared) sl.prepare();
    sl.fpot.NL_UpdateIfNeeded(sl.atoms);
    sl.fpot.compute(sl.atoms);
    for (size_t i = 0; i < sl.fpot.lastE.size(); i++) e_terms[i] = sl.fpot.lastE[i];
    if (grad)
      for (size_t i = 0; i < sl.atoms.size(); i++) {
        grad[3 * i] = sl.atoms[i].grad.x;
        grad[3 * i + 1] = sl.atoms[i].grad.y;
        grad[3 * i + 2] = sl.atoms[i].grad.z;
      }
  });
}

// Parameter override (the harness-subclass route: REBO's pair parameters are
// protected members, REBO.hpp:155-175; the LJ ones private, AIREBO_LJ.hpp).
// theta layout: for p = C-C, C-H, H-H, {Q, A, alpha, B1-3, beta1-3} at 9p and
// {epsilon, sigma} at 27 + 2p; sigma also sets RLJ = [sigma, 2^(1/6) sigma] as
// AIREBO::setupPotential does (AIREBO_LJ.cxx:239-247).
static void theta_rw(Handle* h, double* th, bool write) {
  const int pa[3] = {REBO::C, REBO::C, REBO::H}, pb[3] = {REBO::C, REBO::H, REBO::H};
  for (Potential* pot : h->sl.fpot.pots) {
    if (auto* r = dynamic_cast<REBO*>(pot)) {
      for (int p = 0; p < 3; p++) {
        int a = pa[p], b = pb[p];
        double* v[9] = {&r->Q_[a][b], &r->A_[a][b], &r->alpha_[a][b], &r->B_[0][a][b], &r->B_[1][a][b],
                        &r->B_[2][a][b], &r->beta_[0][a][b], &r->beta_[1][a][b], &r->beta_[2][a][b]};
        double* w[9] = {&r->Q_[b][a], &r->A_[b][a], &r->alpha_[b][a], &r->B_[0][b][a], &r->B_[1][b][a],
                        &r->B_[2][b][a], &r->beta_[0][b][a], &r->beta_[1][b][a], &r->beta_[2][b][a]};
        for (int k = 0; k < 9; k++) {
          if (write) *v[k] = *w[k] = th[9 * p + k];
          else if (!dynamic_cast<ETors*>(pot)) th[9 * p + k] = *v[k];
        }
      }
    }
    if (auto* l = dynamic_cast<AIREBO_LJ*>(pot)) {
      for (int p = 0; p < 3; p++) {
        int a = pa[p], b = pb[p];
        if (write) {
          l->zeta_[a][b] = l->zeta_[b][a] = th[27 + 2 * p];
          l->sigma_[a][b] = l->sigma_[b][a] = th[28 + 2 * p];
          for (int q = 0; q < 2; q++) {
            int x = q ? b : a, y = q ? a : b;
            l->RLJ_[x][y][0] = l->sigma_[x][y];
            l->RLJ_[x][y][1] = l->sigma_[x][y] * std::pow(2.0, 1.0 / 6.0);
          }
        } else {
          th[27 + 2 * p] = l->zeta_[a][b];
          th[28 + 2 * p] = l->sigma_[a][b];
        }
      }
    }
  }
}

int orc_set_params(void* p, const double* theta) {
  std::vector<double> th(theta, theta + 33);
  return guarded([&] { theta_rw(static_cast<Handle*>(p), th.data(), true); });
}

int orc_get_params(void* p, double* theta) {
  return guarded([&] { theta_rw(static_cast<Handle*>(p), theta, false); });
}

// Thermal bath (SimLoop.hpp:118-189): geometry TB_NONE/UNIVERSE/BOX/SPHERE,
// To (K), gamma (1/s); lengths in cm.  zMinOfFreeZone is the selected
// geometry's (box or sphere) own field.
int orc_set_bath(void* p, int geom, double To, double gamma, double zMin, double dBoundary, double zMinOfFreeZone,
                 const double* center, double radius) {
  Handle* h = static_cast<Handle*>(p);
  return guarded([&] {
    if (geom < TB_NONE || geom > TB_SPHERE) throw OracleError("bad thermal bath geometry");
    auto& b = h->sl.bath;
    b.geom = geom;
    b.To = To;
    b.gamma = gamma;
    b.zMin = zMin;
    b.dBoundary = dBoundary;
    b.zMinOfFreeZone = zMinOfFreeZone;
    b.sphereZMinOfFreeZone = zMinOfFreeZone;
    if (center) b.center = V3(center[0], center[1], center[2]);
    b.radius = radius;
  });
}

// Atom::apply_ThermalBath per atom (null: all true).
int orc_set_bath_atoms(void* p, const unsigned char* apply) {
  Handle* h = static_cast<Handle*>(p);
  auto& at = h->sl.atoms;
  for (size_t i = 0; i < at.size(); i++) at[i].apply_tb = apply ? apply[i] != 0 : true;
  return 0;
}

// Energy transferred from the bath so far (erg, SimLoop::check).
int orc_bath_energy(void* p, double* e) {
  *e = static_cast<Handle*>(p)->sl.energyTransferredFromBath;
  return 0;
}

// Neighbour list of chain entry `term`: counts[n], then the concatenated lists
// (ascending index per atom) into idx (capacity cap).  Returns -1 on overflow.
int orc_nl(void* p, int term, int* counts, int* idx, long cap) {
  Handle* h = static_cast<Handle*>(p);
  auto& sl = h->sl;
  int rc = guarded([&] {
    if (!sl.prepared) sl.prepare();
    sl.fpot.NL_UpdateIfNeeded(sl.atoms);
  });
  if (rc) return rc;
  auto& nl = sl.fpot.pots[term]->nl.nl;
  long pos = 0;
  for (size_t i = 0; i < nl.size(); i++) {
    counts[i] = (int)nl[i].size();
    for (int j : nl[i]) {
      if (pos >= cap) return -1;
      idx[pos++] = j;
    }
  }
  return 0;
}

// Run nsteps iterations of the Verlet loop.  fixed_dt > 0 fixes dt (seconds);
// otherwise the reference's adaptive rule is used.  Per step: epot (erg, at
// t+dt), ekin (erg, after the velocity update: synchronised with epot) and the
// dt that was used.  Any output pointer may be null.
int orc_run(void* p, int nsteps, double fixed_dt, double* epot, double* ekin, double* dts,
            double* netforce) {
  Handle* h = static_cast<Handle*>(p);
  return guarded([&] {
    auto& sl = h->sl;
    sl.fixed_dt = fixed_dt > 0;
    sl.fixed_dt_value = fixed_dt;
    for (int s = 0; s < nsteps; s++) {
      double dt_used = sl.fixed_dt ? fixed_dt : sl.dt;
      sl.step();
      if (epot) epot[s] = sl.Epot;
      if (ekin) ekin[s] = sl.Ekin;
      if (dts) dts[s] = dt_used;
      if (netforce) netforce[s] = norm(sl.netForce);
    }
  });
}

// Each atom's own PBC vector and image counters (Atom::PBC, Atom::PBC_count;
// mdtk/Atom.hpp:35-82), set before the first call: prepare() wraps every
// atom by its own PBC (AtomsContainer.cxx:106-123) and depos takes both
// atoms' PBC (Atom.hpp:115-140).  The array box stays the one given to
// orc_create (checkMIC).
int orc_set_atom_pbc(void* p, const double* pbc, const int* pbc_count) {
  Handle* h = static_cast<Handle*>(p);
  return guarded([&] {
    if (h->sl.prepared) throw OracleError("orc_set_atom_pbc: set before the first call");
    auto& at = h->sl.atoms;
    for (size_t i = 0; i < at.size(); i++) {
      at[i].PBC = V3(pbc[3 * i], pbc[3 * i + 1], pbc[3 * i + 2]);
      if (pbc_count)
        for (int d = 0; d < 3; d++) at[i].pbc_count[d] = pbc_count[3 * i + d];
    }
  });
}

int orc_get_state(void* p, double* x, double* v, double* a, int* pbc_count) {
  Handle* h = static_cast<Handle*>(p);
  auto& at = h->sl.atoms;
  for (size_t i = 0; i < at.size(); i++) {
    if (x) { x[3 * i] = at[i].coords.x; x[3 * i + 1] = at[i].coords.y; x[3 * i + 2] = at[i].coords.z; }
    if (v) { v[3 * i] = at[i].V.x; v[3 * i + 1] = at[i].V.y; v[3 * i + 2] = at[i].V.z; }
    if (a) { a[3 * i] = at[i].an.x; a[3 * i + 1] = at[i].an.y; a[3 * i + 2] = at[i].an.z; }
    if (pbc_count)
      for (int d = 0; d < 3; d++) pbc_count[3 * i + d] = at[i].pbc_count[d];
  }
  return 0;
}

// Evaluate one spline table of the oracle at (x,y,z): table ids
// 0 G_C1, 1 G_C2, 2 G_H (x = cos), 3 P_CC, 4 P_CH (x=NH, y=NC),
// 5 pi_rc_CC, 6 pi_rc_CH, 7 pi_rc_HH, 8 T_ij (x=Ni, y=Nj, z=Nconj),
// 9/10 Brenner-I H_CC/H_CH, 11 Brenner-I F, 12-14 the same for Brenner-II.
// which: 0 value, 1..3 partial derivative along x,y,z.
int orc_spline_eval(int table, double x, double y, double z, int which, double* out) {
  static REBO* rebo = nullptr;
  static Brenner* b1 = nullptr;
  static Brenner* b2 = nullptr;
  return guarded([&] {
    if (!rebo) rebo = new REBO();
    if (!b1) b1 = new Brenner(0);
    if (!b2) b2 = new Brenner(1);
    double r = 0;
    switch (table) {
      case 0: r = which ? rebo->gC1.d(x) : rebo->gC1(x); break;
      case 1: r = which ? rebo->gC2.d(x) : rebo->gC2(x); break;
      case 2: r = which ? rebo->gH.d(x) : rebo->gH(x); break;
      case 3: r = rebo->pCC.eval(x, y, which); break;
      case 4: r = rebo->pCH.eval(x, y, which); break;
      case 5: r = rebo->rcCC.eval(x, y, z, which); break;
      case 6: r = rebo->rcCH.eval(x, y, z, which); break;
      case 7: r = rebo->rcHH.eval(x, y, z, which); break;
      case 8: r = rebo->tij.eval(x, y, z, which); break;
      case 9: r = b1->hCC.eval(x, y, which); break;
      case 10: r = b1->hCH.eval(x, y, which); break;
      case 11: r = b1->fF.eval(x, y, z, which); break;
      case 12: r = b2->hCC.eval(x, y, which); break;
      case 13: r = b2->hCH.eval(x, y, which); break;
      case 14: r = b2->fF.eval(x, y, z, which); break;
      default: throw OracleError("unknown table");
    }
    *out = r;
  });
}
// ---------------------------------------------------------------- pieces
// Single functions of the restatement, for checks against the reference's
// own compiled code (oracle/_ref) and against closed forms.

// depos(a1, a2) with per-atom PBC vectors and counters (Atom.hpp:115-140).
int orc_depos(const double* xa, const double* La, const int* ca, const double* xb, const double* Lb, const int* cb,
              double* out) {
  return guarded([&] {
    Atom a, b;
    a.coords = V3(xa[0], xa[1], xa[2]);
    b.coords = V3(xb[0], xb[1], xb[2]);
    a.PBC = V3(La[0], La[1], La[2]);
    b.PBC = V3(Lb[0], Lb[1], Lb[2]);
    for (int d = 0; d < 3; d++) {
      a.pbc_count[d] = ca[d];
      b.pbc_count[d] = cb[d];
    }
    V3 r = depos(a, b);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
  });
}

// Atom::applyPBC (Atom.cxx:93-125) on one atom.
int orc_apply_pbc(double* xyz, const double* L, int* pc) {
  return guarded([&] {
    Atom a;
    a.coords = V3(xyz[0], xyz[1], xyz[2]);
    a.PBC = V3(L[0], L[1], L[2]);
    for (int d = 0; d < 3; d++) a.pbc_count[d] = pc[d];
    a.applyPBC();
    xyz[0] = a.coords.x;
    xyz[1] = a.coords.y;
    xyz[2] = a.coords.z;
    for (int d = 0; d < 3; d++) pc[d] = a.pbc_count[d];
  });
}

// CosTheta(ij, ik, V) for atoms i, j, k (open boundaries): out[0] = cos,
// out[1..9] the gradients added to i, j, k.
int orc_cos_theta(const double* xyz, double V, double* out) {
  return guarded([&] {
    Atom at[3];
    for (int q = 0; q < 3; q++) at[q].coords = V3(xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]);
    Pair ij(at[0], at[1], true), ik(at[0], at[2], true);
    out[0] = CosTheta(ij, ik, V);
    for (int q = 0; q < 3; q++) {
      out[1 + 3 * q] = at[q].grad.x;
      out[2 + 3 * q] = at[q].grad.y;
      out[3 + 3 * q] = at[q].grad.z;
    }
  });
}

// CosDihedral(ij, ik, jl, V) for atoms i, j, k, l: out[0] = cos, out[1..12]
// the gradients added to i, j, k, l.
int orc_cos_dihedral(const double* xyz, double V, double* out) {
  return guarded([&] {
    Atom at[4];
    for (int q = 0; q < 4; q++) at[q].coords = V3(xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]);
    Pair ij(at[0], at[1], true), ik(at[0], at[2], true), jl(at[1], at[3], true);
    out[0] = CosDihedral(ij, ik, jl, V);
    for (int q = 0; q < 4; q++) {
      out[1 + 3 * q] = at[q].grad.x;
      out[2 + 3 * q] = at[q].grad.y;
      out[3 + 3 * q] = at[q].grad.z;
    }
  });
}

// NeighbourList::MovedTooMuch (NeighbourList.cxx:108-133) over disp[n*3].
int orc_moved_too_much(int n, const double* disp, double rc, int* out) {
  return guarded([&] {
    NeighbourList nl;
    nl.Rcutoff = rc;
    nl.init(n);
    for (int i = 0; i < n; i++) nl.displacements[i] = V3(disp[3 * i], disp[3 * i + 1], disp[3 * i + 2]);
    *out = nl.movedTooMuch() ? 1 : 0;
  });
}

// The C_ij path the AIREBO-LJ member stored for (i, j) at the last force call
// (AIREBO_LJ.cxx:289-358): *w the path weight (0 when there is none) and
// nodes[0..hops] the atoms along it (nodes[0] = i); *hops = 0 for none.
int orc_cij_path(void* p, int i, int j, double* w, int* hops, int* nodes) {
  Handle* h = static_cast<Handle*>(p);
  return guarded([&] {
    AIREBO_LJ* lj = nullptr;
    for (auto* q : h->sl.fpot.pots)
      if (auto* c = dynamic_cast<AIREBO_LJ*>(q)) lj = c;
    if (!lj) throw OracleError("no AIREBO-LJ member");
    if (i < 0 || j < 0 || (size_t)i >= lj->CA.size() || (size_t)j >= lj->CA.size())
      throw OracleError("atom index out of range (call energy_grad first)");
    *w = 0.0;
    *hops = 0;
    auto it = lj->CA[i].find((size_t)j);
    if (it == lj->CA[i].end()) return;
    *w = it->second.w;
    const auto& path = it->second.path;
    *hops = (int)path.size();
    for (size_t q = 0; q < path.size(); q++) {
      nodes[q] = (int)path[q].a1->globalIndex;
      nodes[q + 1] = (int)path[q].a2->globalIndex;
    }
  });
}
}
