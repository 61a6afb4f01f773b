// C entry points over the reference's own GSL-free translation units
// (compiled in place from the reference tree by oracle/ref.mk into
// oracle/_ref/libmdtkref.so).
//
// TEST INFRASTRUCTURE ONLY: the CPU tests load this library to check the
// oracle's restatement against the reference code itself where that code
// builds here -- FBZL (pairwise/FBZL.cxx), FBM (pairwise/FBM.cxx), the
// neighbour-list predicate, order and skin rule (potentials/NeighbourList.cxx),
// depos with mixed PBC (Atom.hpp:115-140), Atom::applyPBC (Atom.cxx:93-125),
// CosTheta / CosDihedral and their adjoints (potentials/FGeneral.hpp:136-305),
// FProxy's chain (potentials/FProxy.cxx:45-61) and
// AtomsArray::removeMomentum (AtomsContainer.cxx:268-278).  The spline-based
// classes (REBO, AIREBO, ETors, Brenner, FLJ, Ackland, TightBinding) need GSL
// (SplineAux.cxx:34) and are not built.  Nothing here is part of the product.
#include <mdtk/Atom.hpp>
#include <mdtk/AtomsContainer.hpp>
#include <mdtk/potentials/FGeneral.hpp>
#include <mdtk/potentials/FProxy.hpp>
#include <mdtk/potentials/NeighbourList.hpp>
#include <mdtk/potentials/pairwise/FBM.hpp>
#include <mdtk/potentials/pairwise/FBZL.hpp>

#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace mdtk;

namespace {
std::string g_err;

// A potential that only owns a neighbour list: the handled elements and the
// cutoff a real member would give it (FGeneral.hpp:54-114 is abstract in
// getRcutoff and operator()).
class ListOnly : public FGeneral {
 public:
  Float rc;
  explicit ListOnly(Float r) : rc(r) { nl.Rcutoff = r; }
  Float getRcutoff() const { return rc; }
  Float operator()(AtomsArray&) { return 0.0; }
};

void fill_atoms(AtomsArray& A, int n, const int* elem, const double* xyz, const double* pbc, const int* pbc_count) {
  A.clear();
  for (int i = 0; i < n; i++) {
    Atom a(ElementID(elem[i]), Vector3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    a.PBC = Vector3D(pbc[0], pbc[1], pbc[2]);
    if (pbc_count) {
      a.PBC_count.x = pbc_count[3 * i];
      a.PBC_count.y = pbc_count[3 * i + 1];
      a.PBC_count.z = pbc_count[3 * i + 2];
    }
    A.push_back(a);
  }
  A.prepareForSimulatation();
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "exception in the reference code";
  }
  return 1;
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// depos(a1, a2) of two atoms with their own PBC vectors and PBC_count
// (Atom.hpp:115-140).
int ref_depos(const double* xa, const double* La, const int* ca, const double* xb, const double* Lb, const int* cb,
              double* out) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    Atom a(C_EL, Vector3D(xa[0], xa[1], xa[2])), b(C_EL, Vector3D(xb[0], xb[1], xb[2]));
    a.PBC = Vector3D(La[0], La[1], La[2]);
    b.PBC = Vector3D(Lb[0], Lb[1], Lb[2]);
    a.PBC_count.x = ca[0]; a.PBC_count.y = ca[1]; a.PBC_count.z = ca[2];
    b.PBC_count.x = cb[0]; b.PBC_count.y = cb[1]; b.PBC_count.z = cb[2];
    Vector3D r = depos(a, b);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
  });
}

// Atom::applyPBC (Atom.cxx:93-125): wrapped coordinates and updated PBC_count.
int ref_apply_pbc(double* xyz, const double* L, int* pc) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    Atom a(C_EL, Vector3D(xyz[0], xyz[1], xyz[2]));
    a.PBC = Vector3D(L[0], L[1], L[2]);
    a.PBC_count.x = pc[0]; a.PBC_count.y = pc[1]; a.PBC_count.z = pc[2];
    a.applyPBC();
    xyz[0] = a.coords.x; xyz[1] = a.coords.y; xyz[2] = a.coords.z;
    pc[0] = a.PBC_count.x; pc[1] = a.PBC_count.y; pc[2] = a.PBC_count.z;
  });
}

// CosTheta(ij, ik, V) (FGeneral.hpp:138-185) for atoms xyz[0..2] = i, j, k:
// out[0] = cos, out[1..9] = the gradients it adds to i, j, k.
int ref_cos_theta(const double* xyz, double V, double* out) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    ListOnly f(2.0 * Ao);
    Atom at[3];
    for (int q = 0; q < 3; q++) {
      at[q] = Atom(C_EL, Vector3D(xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]));
      at[q].grad = Vector3D(0, 0, 0);
    }
    AtomsPair ij(at[0], at[1], true), ik(at[0], at[2], true);
    out[0] = f.CosTheta(ij, ik, V);
    for (int q = 0; q < 3; q++) {
      out[1 + 3 * q] = at[q].grad.x; out[2 + 3 * q] = at[q].grad.y; out[3 + 3 * q] = at[q].grad.z;
    }
  });
}

// CosDihedral(ij, ik, jl, V) (FGeneral.hpp:201-305) for atoms i, j, k, l:
// out[0] = cos, out[1..12] = the gradients it adds to i, j, k, l.
int ref_cos_dihedral(const double* xyz, double V, double* out) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    ListOnly f(2.0 * Ao);
    Atom at[4];
    for (int q = 0; q < 4; q++) {
      at[q] = Atom(C_EL, Vector3D(xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]));
      at[q].grad = Vector3D(0, 0, 0);
    }
    AtomsPair ij(at[0], at[1], true), ik(at[0], at[2], true), jl(at[1], at[3], true);
    out[0] = f.CosDihedral(ij, ik, jl, V);
    for (int q = 0; q < 4; q++) {
      out[1 + 3 * q] = at[q].grad.x; out[2 + 3 * q] = at[q].grad.y; out[3 + 3 * q] = at[q].grad.z;
    }
  });
}

// Energy (erg) and gradient (erg/cm) of a chain of the reference's buildable
// pairwise members through FProxy::operator() (FProxy.cxx:45-61): mask bit 1
// FBZL with the switch r0..r1 (common.h:46 uses 5.0/5.5 A), bit 128 FBM
// (default Rcutoff 3.1/3.5 A, FPairwise.hpp:34).  pbc: the box (NO_PBC for
// an open axis); lists built once at these positions.
int ref_pair_chain(int n, const int* elem, const double* xyz, const double* pbc, int mask, double r0, double r1,
                   double* e_terms, double* grad) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    AtomsArray A;
    fill_atoms(A, n, elem, xyz, pbc, nullptr);
    FProxy chain;
    if (mask & 1) chain.addPotential(new FBZL(Rcutoff(r0, r1)));
    if (mask & 128) chain.addPotential(new FBM());
    chain.NL_init(A);
    chain.NL_UpdateIfNeeded(A);
    for (size_t i = 0; i < A.size(); i++) A[i].grad = Vector3D(0, 0, 0);
    for (size_t t = 0; t < chain.potentials.size(); t++) e_terms[t] = (*chain.potentials[t])(A);
    for (int i = 0; i < n; i++) {
      grad[3 * i] = A[i].grad.x; grad[3 * i + 1] = A[i].grad.y; grad[3 * i + 2] = A[i].grad.z;
    }
    chain.freePotentials();
  });
}

// NeighbourList::Update (NeighbourList.cxx:31-106) for one list of cutoff rc
// over the atoms whose element is in handled[0..nh): counts[n] and the rows
// concatenated into idx[cap] in the reference's order.
int ref_neighbour_list(int n, const int* elem, const double* xyz, const double* pbc, double rc, const int* handled,
                       int nh, int* counts, int* idx, long cap) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    AtomsArray A;
    fill_atoms(A, n, elem, xyz, pbc, nullptr);
    ListOnly f(rc);
    for (int q = 0; q < nh; q++) f.handledElements.insert(ElementID(handled[q]));
    f.NL_init(A);
    f.NL_UpdateIfNeeded(A);
    long pos = 0;
    for (int i = 0; i < n; i++) {
      AtomRefsContainer& row = f.NL(A[i]);
      counts[i] = (int)row.size();
      for (size_t q = 0; q < row.size(); q++) {
        if (pos >= cap) throw std::runtime_error("ref_neighbour_list: capacity");
        idx[pos++] = (int)row[q]->globalIndex;
      }
    }
  });
}

// NeighbourList::MovedTooMuch (NeighbourList.cxx:108-133) over the
// accumulated displacements disp[n*3]: 1 when a rebuild is due.
int ref_moved_too_much(int n, const double* disp, double rc, int* out) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    AtomsArray A;
    std::vector<int> el(n, C_EL);
    std::vector<double> x(3 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++) x[3 * i] = i * 1e-7;
    double L[3] = {NO_PBC.x, NO_PBC.y, NO_PBC.z};
    fill_atoms(A, n, el.data(), x.data(), L, nullptr);
    ListOnly f(rc);
    f.NL_init(A);
    for (int i = 0; i < n; i++) f.incDisplacement(A[i], Vector3D(disp[3 * i], disp[3 * i + 1], disp[3 * i + 2]));
    *out = f.nl.MovedTooMuch(A) ? 1 : 0;
  });
}

// AtomsArray::removeMomentum (AtomsContainer.cxx:268-278) on velocities
// vel[n*3] (cm/s) of atoms elem[n] with fixed flags (may be null).
int ref_remove_momentum(int n, const int* elem, const unsigned char* fixed, double* vel) {
  yaatk::verboseTrace = false;
  return guarded([&] {
    AtomsArray A;
    std::vector<double> x(3 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++) x[3 * i] = i * 1e-7;
    double L[3] = {NO_PBC.x, NO_PBC.y, NO_PBC.z};
    fill_atoms(A, n, elem, x.data(), L, nullptr);
    for (int i = 0; i < n; i++) {
      A[i].V = Vector3D(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
      if (fixed && fixed[i]) A[i].fix();
    }
    A.removeMomentum();
    for (int i = 0; i < n; i++) {
      vel[3 * i] = A[i].V.x; vel[3 * i + 1] = A[i].V.y; vel[3 * i + 2] = A[i].V.z;
    }
  });
}

}  // extern "C"
