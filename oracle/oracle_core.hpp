// Oracle: CPU restatement of the MDTK 0.8.0 force-and-integration path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this code, and only as the checker
// (or as the timed CPU baseline).  The GPU product never links or calls it.
//
// It restates, in this project's own code, the reference's algorithm and its
// evaluation structure: AoS atoms, per-potential O(N^2) Verlet-skin neighbour
// lists, and the "seeded" reverse-mode scheme in which every sub-term f(pair,V)
// returns its value and, when V != 0, adds V*df/dx into Atom::grad.  Each
// function cites the reference file:line it follows.
//
// Documented deviations (see the per-function comments):
//  * Splines.  The reference solves a monomial system in absolute coordinates
//    with GSL's LU (mdtk/SplineAux.cxx:36-93) and evaluates the monomial sum in
//    absolute coordinates (mdtk/SplineAux.hpp:80-94,230-241).  GSL is not
//    installed, so the reference coefficient tables cannot be reproduced bit
//    for bit.  The oracle builds the same equations (same unknown ordering,
//    same conditions) in cell-local coordinates x-x0 and solves them with its
//    own partial-pivot LU; this is the same unique Hermite interpolant,
//    evaluated without the absolute-coordinate cancellation.
//  * The CosTheta / CosDihedral adjoints are derived here in vector form; they
//    are the gradients of the same functions as mdtk/potentials/FGeneral.hpp:
//    136-305, equal up to rounding.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// CGS constants, mdtk/consts.cxx:28-38, mdtk/consts.hpp:96
constexpr double Ao = 1e-8;
constexpr double eV = 1.6021E-012;
constexpr double amu = 1.6605655e-24;
constexpr double e_charge = 4.8e-10;
constexpr double kb = 8.617343E-5 * eV;
constexpr double NO_PBC_L = 10000.0 * Ao;

// element ids, mdtk/consts.hpp:46-56
enum Elem { H_EL = 1, C_EL = 12, Cu_EL = 64, Ag_EL = 108, Au_EL = 197, Ar_EL = 40, Xe_EL = 131 };

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
};
inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(const V3& a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(const V3& a, double m) { return V3(a.x * m, a.y * m, a.z * m); }
inline V3 operator*(double m, const V3& a) { return V3(a.x * m, a.y * m, a.z * m); }
inline V3 operator/(const V3& a, double m) { return V3(a.x / m, a.y / m, a.z / m); }
inline void operator+=(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double norm(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline double norm2(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double SQR(double x) { return x * x; }

struct OracleError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
#define ORC_REQUIRE(c)                                                   \
  do {                                                                   \
    if (!(c)) throw ::orc::OracleError(std::string("REQUIRE failed: ") + #c); \
  } while (0)

// mdtk/Atom.hpp:32-96 (fields used on the hot path)
struct Atom {
  int ID = C_EL;
  double Z = 0, M = 0;
  V3 coords, V, an, grad;
  int pbc_count[3] = {0, 0, 0};
  bool fixed = false;
  bool apply_tb = true;  // Atom::apply_ThermalBath (Atom.hpp:52)
  V3 an_no_tb;           // acceleration without the bath's force
  V3 PBC{NO_PBC_L, NO_PBC_L, NO_PBC_L};
  size_t globalIndex = 0;
  void setAttributesByElementID();  // mdtk/Atom.hpp:98-113
  void applyPBC();                  // mdtk/Atom.cxx:93-125
};

V3 depos(const Atom& a1, const Atom& a2);  // mdtk/Atom.hpp:115-140

// Virial of the gradient pushes (no reference counterpart): while g_vir_on,
// every primitive that adds V*d/dx into Atom::grad (Pair::r, Pair::r_alter,
// CosTheta, CosDihedral) also adds -r (x) v for each of its pair vectors r
// and the push v the pair carries; components xx, yy, zz, xy, xz, yz (erg).
extern bool g_vir_on;
extern double g_vir[6];
inline void vir_push(const V3& r, const V3& v) {
  g_vir[0] -= r.x * v.x;
  g_vir[1] -= r.y * v.y;
  g_vir[2] -= r.z * v.z;
  g_vir[3] -= r.x * v.y;
  g_vir[4] -= r.x * v.z;
  g_vir[5] -= r.y * v.z;
}

// Test switch: when true, the b* pair's r_alter gets a zero derivative (the
// exact derivative of its constant value) instead of the reference's
// unit-vector derivative.  Default false (reference behaviour).
extern bool g_alter_exact;

// mdtk/potentials/AtomsPair.hpp:35-158
struct Pair {
  Atom* a1;
  Atom* a2;
  V3 rv;
  double r_ = 0, r2_ = 0;
  V3 u;  // rv / r_
  double f_ = 0, df_ = 0;
  double ralt_ = 0;
  bool alt = false;
  Pair(Atom& i, Atom& j, bool init);
  Pair(Atom& i, Atom& j, double R1, double R2, double alter = 0.0);
  Pair neg() const;
  V3 dr(const Atom& d) const {
    if (a1 == &d) return u;
    if (a2 == &d) return -u;
    return V3();
  }
  V3 dr_alter(const Atom& d) const;
  double r(double V = 0.0) {
    if (V != 0.0) {
      a1->grad += dr(*a1) * V;
      a2->grad += dr(*a2) * V;
      if (g_vir_on) vir_push(rv, u * V);
    }
    return r_;
  }
  double r_alter(double V = 0.0);
  double f(double V = 0.0) {
    if (df_ != 0.0) r(V * df_);
    return f_;
  }
};

double CosTheta(Pair& ij, Pair& ik, double V = 0.0);          // FGeneral.hpp:138-185
double SinTheta(Pair& ij, Pair& ik);                          // FGeneral.hpp:187-199
double CosDihedral(Pair& ij, Pair& ik, Pair& jl, double V = 0.0);  // FGeneral.hpp:201-305

using AtomsArray = std::vector<Atom>;

// mdtk/potentials/NeighbourList.{hpp,cxx}
struct NeighbourList {
  double Rcutoff = 0;
  bool updateRequested = true;
  std::vector<std::vector<int>> nl;
  std::vector<V3> displacements;
  void init(size_t n) {
    nl.assign(n, {});
    displacements.assign(n, V3());
    updateRequested = true;
  }
  bool movedTooMuch() const;  // NeighbourList.cxx:108-133
};

// ---------------------------------------------------------------- splines
// One cell of a tensor-product Hermite interpolant, built as in
// mdtk/SplineAux.hpp:133-206 (monomials enumerated with the first coordinate's
// power fastest) and mdtk/Spline{5n,2D,3D}.cxx (value/derivative equations),
// but in cell-local coordinates.
struct LocalPoly {
  int dim = 0, order = 3;
  std::vector<int> pw;      // nterms*dim powers
  std::vector<double> c;    // coefficients (0 => default-constructed spline, value 0)
  double x0[3] = {0, 0, 0};
  bool built = false;
  void setup(int dim, int order);
  double eval(const double* x, int dx = -1) const;  // value, or d/dx_dx
};

struct Spline5nO {  // mdtk/Spline5n.cxx:34-105
  LocalPoly p;
  double xa = 0, xb = 0;
  void init(double x1, double y1, double dy1, double d2y1, double x2, double y2, double dy2, double d2y2);
  double val(double x) const { return p.eval(&x); }
  double der(double x) const { return p.eval(&x, 0); }
};

struct Spline2DO {  // mdtk/Spline2D.cxx:35-125
  LocalPoly p;
  void init(double x0, double y0, const double z[2][2], const double dzdx[2][2], const double dzdy[2][2]);
};

struct Spline3DO {  // mdtk/Spline3D.cxx:34-163
  LocalPoly p;
  void init(double x0, double y0, double z0, const double v[2][2][2], const double dvdx[2][2][2],
            const double dvdy[2][2][2], const double dvdz[2][2][2]);
};

// AIREBO/cubici.hpp:39-80 G(cos) piecewise quintic
struct FuncGO {
  Spline5nO sp[4];
  int count = 0;
  void init(const double* x, const double* y, const double* dy, const double* d2y, int npts);
  int cell(double c) const;
  double operator()(double c) const { return sp[cell(c)].val(c); }
  double d(double c) const { return sp[cell(c)].der(c); }
};

// bicubic P_CC/P_CH (AIREBO/cubici.hpp:80-222) and Brenner H_CC/H_CH
struct Func2O {
  Spline2DO sp[4][4];
  double intarg[4][4], dH_intarg[4][4], dC_intarg[4][4];
  bool use_intarg = true;
  void init(const double node[5][5], const double dx[5][5], const double dy[5][5], bool intarg_shortcut);
  double eval(double h, double c, int which) const;  // which: 0 val, 1 d/dh, 2 d/dc
};

// tricubic pi_rc / Tij (AIREBO/cubici.hpp:123-369) and Brenner F (kmax=4, no intarg)
struct Func3O {
  int ni = 5, nj = 5, nk = 16;  // node extents
  int kclamp = 15;
  bool use_intarg = true;
  std::vector<Spline3DO> sp;     // (ni-1)*(nj-1)*(nk) cells; cells with k >= built_k are default
  std::vector<double> iav, dia, dja, dka;  // fill_intarg caches
  int built_k = 15;
  void init(int ni, int nj, int nk, int built_k, int kclamp, bool intarg_shortcut,
            const std::vector<double>& v, const std::vector<double>& vi, const std::vector<double>& vj,
            const std::vector<double>& vk);
  double eval(double i, double j, double k, int which) const;  // 0 val, 1 di, 2 dj, 3 dk
};

// ---------------------------------------------------------------- potentials
struct Potential {
  NeighbourList nl;
  std::vector<int> handled;  // element ids
  bool isHandled(const Atom& a) const {
    for (int e : handled)
      if (e == a.ID) return true;
    return false;
  }
  virtual double getRcutoff() const = 0;
  virtual double compute(AtomsArray& atoms) = 0;
  virtual ~Potential() {}
  std::vector<int>& NL(const Atom& a) { return nl.nl[a.globalIndex]; }
};

struct REBO : Potential {  // mdtk/potentials/manybody/AIREBO/REBO.{hpp,cxx}
  enum { C = 0, H = 1 };
  double Q_[2][2], alpha_[2][2], A_[2][2], B_[3][2][2], beta_[3][2][2], rho_[2][2];
  double R_[2][2][2], Rp_[2][2][2], Ntot_[2];
  FuncGO gC1, gC2, gH;
  Func2O pCC, pCH;
  Func3O rcCC, rcCH, rcHH, tij;
  std::vector<double> Nts, NCs, NHs;
  std::vector<std::vector<int>> Nts2t, NCs2t, NHs2t;
  AtomsArray* gl = nullptr;
  REBO();
  double getRcutoff() const override { return std::max(R_[C][C][1], std::max(R_[C][H][1], R_[H][H][1])); }
  int e2i(const Atom& a) const {
    if (a.ID == H_EL) return H;
    if (a.ID == C_EL) return C;
    throw OracleError("e2i() : unknown element");
  }
  double R(int i, const Atom& a, const Atom& b) const { return R_[e2i(a)][e2i(b)][i]; }
  bool probablyAreNeighbours(const Atom& a, const Atom& b) const {
    return !(norm2(depos(a, b)) > SQR(R(1, a, b)));
  }
  double compute(AtomsArray& atoms) override;
  void countNeighbours(AtomsArray& atoms);
  double Sprime(double a, double mn, double mx) const;
  double dSprime(double a, double mn, double mx) const;
  double fprime(Pair& ij, double V = 0.0);
  double VR(Pair& ij, double V = 0.0);
  double VA(Pair& ij, double V = 0.0);
  double Baver(Pair& ij, double V = 0.0);
  double ExpTerm(Pair& ij, Pair& ik, double V = 0.0);
  double D(Pair& ij, double V = 0.0);
  double p_sigma_pi(Pair& ij, double V = 0.0);
  double G(Pair& ij, Pair& ik, double V = 0.0);
  void dfThem(const std::vector<std::vector<int>>& patoms, Pair& ij, double V);
  double Nt(Pair& ij, double V = 0.0);
  double NH(Pair& ij, double V = 0.0);
  double NC(Pair& ij, double V = 0.0);
  double NconjSum(Pair& ij, bool second, double V = 0.0);
  double Nconj(Pair& ij, double V = 0.0);
  double P(Pair& ij, double V = 0.0);
  double pi_rc(Pair& ij, double V = 0.0);
  double pi_dh_sum(Pair& ij, double V = 0.0);
  double pi_dh_Tij(Pair& ij, double V = 0.0);
  double pi_dh(Pair& ij, double V = 0.0);
};

struct AIREBO_LJ : Potential {  // mdtk/potentials/manybody/AIREBO/AIREBO_LJ.{hpp,cxx}
  REBO& rebo;
  double R_[2][2][2], RLJ_[2][2][2], b_[2][2][2], sigma_[2][2], zeta_[2][2];
  struct CEl {
    double w = 0.0;
    std::vector<Pair> path;
  };
  std::vector<std::map<size_t, CEl>> CA;
  explicit AIREBO_LJ(REBO& r);
  double getRcutoff() const override { return std::max(R_[0][0][1], std::max(R_[0][1][1], R_[1][1][1])); }
  double R(int i, const Atom& a, const Atom& b) const { return R_[rebo.e2i(a)][rebo.e2i(b)][i]; }
  bool probablyAreNeighbours(const Atom& a, const Atom& b) const {
    return !(norm2(depos(a, b)) > SQR(R(1, a, b)));
  }
  double compute(AtomsArray& atoms) override;
  double S(double a, double mn, double mx) const;
  double dS(double a, double mn, double mx) const;
  double StrStb(Pair& ij, double V = 0.0);
  double VLJ(Pair& ij, double V = 0.0);
  double BijAsterix(Pair& ij, double V = 0.0);
  void fill_Cij(AtomsArray& atoms);
  double Cij(Pair& ij, double V = 0.0);
};

struct ETors : REBO {  // mdtk/potentials/manybody/AIREBO/AIREBO_ETors.{hpp,cxx}
  double zetaCC_[2][2];
  bool compat = true;  // reproduce the reference's torsion derivative (missing factor 5)
  ETors();
  double compute(AtomsArray& atoms) override;
  double Vtors(Pair& ij, Pair& ik, Pair& jl, double V);
};

struct Brenner : Potential {  // mdtk/potentials/manybody/Brenner/Brenner.{hpp,cxx}
  enum { C = 0, H = 1 };
  int paramSet;
  double Re_[2][2], De_[2][2], beta_[2][2], S_[2][2], delta_[2][2], R_[2][2][2], alpha_[2][2][2];
  double ao_, co_, do_, G_HH_;
  Func2O hCC, hCH;
  Func3O fF;
  AtomsArray* gl_ = nullptr;
  explicit Brenner(int paramSet);
  double getRcutoff() const override { return std::max(R_[C][C][1], std::max(R_[C][H][1], R_[H][H][1])); }
  int e2i(const Atom& a) const {
    if (a.ID == H_EL) return H;
    if (a.ID == C_EL) return C;
    throw OracleError("e2i() : unknown element");
  }
  double R(int i, const Atom& a, const Atom& b) const { return R_[e2i(a)][e2i(b)][i]; }
  bool probablyAreNeighbours(const Atom& a, const Atom& b) const {
    return !(norm2(depos(a, b)) > SQR(R(1, a, b)));
  }
  double compute(AtomsArray& atoms) override;
  double VR(Pair& ij, double V = 0.0);
  double VA(Pair& ij, double V = 0.0);
  double Baver(Pair& ij, double V = 0.0);
  double ExpTerm(Pair& ij, Pair& ik, double V = 0.0);
  double D(Pair& ij, double V = 0.0);
  double B(Pair& ij, double V = 0.0);
  double G(Pair& ij, Pair& ik, double V = 0.0);
  double Nt(Pair& ij, double V = 0.0);
  double NH(Pair& ij, double V = 0.0);
  double NC(Pair& ij, double V = 0.0);
  double F(double x) const;
  double dF(double x) const;
  double Nconj(Pair& ij, double V = 0.0);
  double HN(Pair& ij, double V = 0.0);
  double FN(Pair& ij, double V = 0.0);
};

struct FBZL : Potential {  // mdtk/potentials/pairwise/FBZL.cxx
  double R0, R1, AB_ = 0.53e-8;
  FBZL(double r0, double r1);
  double getRcutoff() const override { return R1; }
  bool isIon(int id) const { return id == Ar_EL || id == Xe_EL; }
  bool isHandledPair(const Atom& a, const Atom& b) const;
  double F11(Pair& ij, double V = 0.0);
  double compute(AtomsArray& atoms) override;
};

// mdtk/potentials/pairwise/FLJ.{hpp,cxx}: hydrocarbon-metal pairs {H,C}x{Cu,Au};
// ZBL below 1.2 A, a cubic bridge (value and slope of both ends) to 1.9 A,
// Lorentz-Berthelot LJ times the cosine cutoff beyond.
struct FLJ : Potential {
  enum { Cu = 0, H = 1, C = 2, Au = 3, ECOUNT = 4 };
  double R0, R1;
  double sigma_[ECOUNT][ECOUNT], epsilon_[ECOUNT][ECOUNT];
  double x_[2] = {1.2 * Ao, 1.9 * Ao};
  double v_[ECOUNT][ECOUNT][2], d_[ECOUNT][ECOUNT][2];  // bridge ends: values and slopes
  FLJ(double r0, double r1);
  double getRcutoff() const override { return R1; }
  int e2i(int id) const;
  bool isHandledPair(const Atom& a, const Atom& b) const;
  double VLJ(Pair& ij);
  double compute(AtomsArray& atoms) override;
};

// mdtk/potentials/manybody/Ackland/Ackland.{hpp,cxx}: the Ackland-Tichy-
// Vitek-Finnis EAM for Cu, Ag, Au.  E = sum_i -sqrt(rho_i) + sum_{i<j} V(r),
// rho_i = sum_j phi(r_ij) (geometric mean of the like-pair phi for unlike
// pairs), V a sum of cubic knots with ZBL below 1.0 A and a cubic bridge
// (value and slope of both ends) to 1.4 A.
struct Ackland : Potential {
  enum { Cu = 0, Ag = 1, Au = 2, ECOUNT = 3 };
  double rk_[ECOUNT][ECOUNT][7] = {}, ak_[ECOUNT][ECOUNT][7] = {};
  double Rk_[ECOUNT][ECOUNT][3] = {}, Ak_[ECOUNT][ECOUNT][3] = {};
  double x_[2] = {1.0 * Ao, 1.4 * Ao};
  double v_[ECOUNT][ECOUNT][2], d_[ECOUNT][ECOUNT][2];  // bridge ends: values and slopes
  Ackland();
  double getRcutoff() const override {  // Ackland.hpp:58
    return std::max(rk_[Au][Au][1], std::max(rk_[Ag][Ag][1], rk_[Cu][Cu][1]));
  }
  int e2i(int id) const;
  double knots(int a, int b, double r, double& der) const;  // the rk/ak cubic sum
  double Phi(Pair& ij);
  double PhiCap(int a, int b, double r) const;
  double dPhiCap(int a, int b, double r) const;
  double g(Pair& ij, double V = 0.0);
  double rho(Atom& a, AtomsArray& atoms, double V = 0.0);
  double F(Atom& a, AtomsArray& atoms);
  double compute(AtomsArray& atoms) override;
};

// mdtk/potentials/pairwise/FBM.{hpp,cxx}: Born-Mayer A exp(-b r) for Cu-Cu
// and Cu-Ar, times the cosine cutoff (R0, R1); not in mdtrajsim's chain.
struct FBM : Potential {
  double R0, R1;
  FBM(double r0, double r1);
  double getRcutoff() const override { return R1; }
  bool isHandledPair(const Atom& a, const Atom& b) const;
  void params(const Atom& a, const Atom& b, double& A, double& B) const;
  double F11(Pair& ij, double V = 0.0);
  double compute(AtomsArray& atoms) override;
};

// mdtk/potentials/manybody/TightBinding/TightBinding.{hpp,cxx}: Cu second-
// moment tight binding, E = sum_i -c sqrt(rho_i) + sum_{i<j} V, rho_i =
// sum_j f(r) exp(-beta r), V = Born-Mayer below 1.0 A, a cubic bridge to
// 1.2 A, f(r) Phi0 exp(-alpha r) beyond; f the cosine cutoff 5.0-5.5 A.
// Not in mdtrajsim's chain.
struct TightBinding : Potential {
  double alpha_, beta_, c_, Phi0_, R_[2], BM_A, BM_B;
  double x_[2] = {1.0 * Ao, 1.2 * Ao};
  double v_[2], d_[2];  // bridge ends: values and slopes
  TightBinding();
  double getRcutoff() const override { return R_[1]; }
  bool near(const Atom& a, const Atom& b) const { return !(norm2(depos(a, b)) > SQR(R_[1])); }
  double Phi(Pair& ij);
  double g(Pair& ij, double V = 0.0);
  double rho(Atom& a, AtomsArray& atoms, double V = 0.0);
  double F(Atom& a, AtomsArray& atoms);
  double compute(AtomsArray& atoms) override;
};

// mdtk/potentials/FProxy.{hpp,cxx} + the NL update in NeighbourList.cxx:31-106
struct Proxy {
  std::vector<Potential*> pots;
  std::vector<double> lastE;
  ~Proxy() {
    for (auto* p : pots) delete p;
  }
  double compute(AtomsArray& atoms);
  void NL_init(AtomsArray& atoms) {
    for (auto* p : pots) p->nl.init(atoms.size());
  }
  void NL_checkRequestUpdate() {
    for (auto* p : pots) p->nl.updateRequested = p->nl.movedTooMuch();
  }
  void NL_UpdateIfNeeded(AtomsArray& atoms);
  void incDisplacement(size_t i, const V3& d) {
    for (auto* p : pots) p->nl.displacements[i] += d;
  }
  bool hasNeighbors(const Atom& a) {
    for (auto* p : pots)
      if (!p->NL(a).empty()) return true;
    return false;
  }
  double getRcutoff() const {
    double m = pots[0]->getRcutoff();
    for (auto* p : pots) m = std::max(m, p->getRcutoff());
    return m;
  }
};

void nl_update(AtomsArray& atoms, std::vector<Potential*>& pots);

// The thermal bath of SimLoop (SimLoop.hpp:118-189): friction
// -V M gamma (1 - sqrt(To/T)) on the atoms inside its geometry, T the kinetic
// temperature of those atoms at the start of the step.
enum { TB_NONE = 0, TB_UNIVERSE = 1, TB_BOX = 2, TB_SPHERE = 3 };
struct ThermalBath {
  int geom = TB_NONE;
  double To = 0.0, gamma = 1.0e13;                                          // SimLoop.hpp:118-127
  double zMin = 1000000.0 * Ao, dBoundary = 0.0, zMinOfFreeZone = -5.0 * Ao;  // :128-152
  V3 center;                                                                // :153-171
  double radius = 1000000.0 * Ao;
  double sphereZMinOfFreeZone = -5.0 * Ao;
  bool inside(const Atom& a) const;  // SimLoop.hpp:139-151,164-170,174-185
};

// mdtk/SimLoop.cxx:225-474 (no file output)
struct SimLoop {
  AtomsArray atoms;
  V3 arrayPBC{NO_PBC_L, NO_PBC_L, NO_PBC_L};  // AtomsArray::arrayPBC (AtomsContainer.hpp:34), checkMIC's box
  Proxy fpot;
  ThermalBath bath;
  double energyTransferredFromBath = 0.0;  // SimLoop.hpp:65
  double dt = 1e-20, simTime = 0;  // SimLoop.cxx:50
  bool fixed_dt = false;
  double fixed_dt_value = 0;
  long iteration = 0;
  double Epot = 0, Ekin = 0;
  V3 netForce;
  bool prepared = false;
  void prepare();
  void step();  // one iteration of the while loop
  double energyKin() const;
  bool bathApplies(const Atom& a) const { return a.apply_tb && bath.inside(a); }  // SimLoop.hpp:186-190
  double bathTemperature() const;  // actualTemperatureOfThermalBath, SimLoop.cxx:516-535
};

}  // namespace orc
