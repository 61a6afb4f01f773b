// Oracle potentials: REBO, AIREBO-LJ, ETors, Brenner, FBZL and the
// velocity-Verlet step.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include <algorithm>

#include "oracle_core.hpp"

namespace orc {

// =================================================================== tables
// Node data transcribed from mdtk/potentials/manybody/AIREBO/cubici.cxx.
namespace {
struct Grid3 {
  int ni, nj, nk;
  std::vector<double> v, vi, vj, vk;
  Grid3(int a, int b, int c) : ni(a), nj(b), nk(c), v(a * b * c, 0.0), vi(v), vj(v), vk(v) {}
  double& at(std::vector<double>& x, int i, int j, int k) { return x[(i * nj + j) * nk + k]; }
  void sym(int i, int j, int k, double val) { at(v, i, j, k) = at(v, j, i, k) = val; }  // FCCfillSym
  void symI(int i, int j, int k, double val) { at(vi, i, j, k) = at(vi, j, i, k) = val; }
  void symJ(int i, int j, int k, double val) { at(vj, i, j, k) = at(vj, j, i, k) = val; }
  void symK(int i, int j, int k, double val) { at(vk, i, j, k) = at(vk, j, i, k) = val; }
  void up(int ib, int jb, int kb) {  // FCCfillUp, cubici.cxx:292-298
    for (int i = 0; i < ni; i++)
      for (int j = 0; j < nj; j++)
        if ((i == ib && j == jb) || (i == jb && j == ib))
          for (int k = kb + 1; k < nk; k++) at(v, i, j, k) = at(v, i, j, kb);
  }
};

void build_pi_rc_CC(Grid3& g) {  // cubici.cxx:367-420
  g.sym(0, 0, 3, 0.004959); g.up(0, 0, 3);
  g.sym(1, 0, 1, 0.021694);
  g.sym(1, 0, 2, 0.004959); g.up(1, 0, 2);
  g.sym(1, 1, 1, 0.052500);
  g.sym(1, 1, 2, -0.002089);
  g.sym(1, 1, 3, -0.008043); g.up(1, 1, 3);
  g.sym(2, 0, 1, 0.024699);
  g.sym(2, 0, 2, -0.005971);
  g.sym(2, 0, 3, 0.004959); g.up(2, 0, 3);
  g.sym(2, 1, 1, 0.004825); g.symI(2, 1, 1, -0.026250);
  g.sym(2, 1, 2, 0.015000);
  g.sym(2, 1, 3, -0.010000);
  g.sym(2, 1, 4, -0.011689); g.symK(2, 1, 4, -0.010022);
  g.sym(2, 1, 5, -0.013378); g.symI(2, 1, 5, -0.027188); g.symK(2, 1, 5, -0.010022);
  g.sym(2, 1, 6, -0.015067); g.symI(2, 1, 6, -0.027188);
  g.sym(2, 1, 7, -0.015067); g.symI(2, 1, 7, -0.027188); g.up(2, 1, 7);
  g.sym(2, 2, 1, 0.047225);
  g.sym(2, 2, 2, 0.011000);
  g.sym(2, 2, 3, 0.019853);
  g.sym(2, 2, 4, 0.016544); g.symK(2, 2, 4, -0.003309);
  g.sym(2, 2, 5, 0.013235); g.symK(2, 2, 5, -0.003309);
  g.sym(2, 2, 6, 0.009926); g.symK(2, 2, 6, -0.003309);
  g.sym(2, 2, 7, 0.006618); g.symK(2, 2, 7, -0.003309);
  g.sym(2, 2, 8, 0.003309); g.symK(2, 2, 8, -0.003309);
  g.sym(3, 0, 1, -0.099899);
  g.sym(3, 0, 2, -0.099899);
  g.sym(3, 0, 3, 0.004959); g.up(3, 0, 3);
  g.sym(3, 1, 2, -0.062418); g.symJ(3, 1, 2, 0.037545);
  g.sym(3, 1, 3, -0.062418); g.up(3, 1, 3);
  g.sym(3, 2, 1, -0.022355);
  g.sym(3, 2, 2, -0.022355); g.symJ(3, 2, 2, 0.062418); g.up(3, 2, 2);
}

void build_pi_rc_CH(Grid3& g) {  // cubici.cxx:425-451
  g.sym(1, 1, 1, -0.050000);
  g.sym(1, 1, 2, -0.050000);
  g.sym(1, 1, 3, -0.300000);
  g.sym(1, 1, 4, -0.050000); g.up(1, 1, 4);
  g.sym(2, 0, 5, -0.004524); g.up(2, 0, 5);
  g.sym(2, 1, 2, -0.250000);
  g.sym(2, 1, 3, -0.250000);
  g.sym(3, 1, 1, -0.100000);
  g.sym(3, 1, 2, -0.125000);
  g.sym(3, 1, 3, -0.125000);
  g.sym(3, 1, 4, -0.100000); g.up(3, 1, 4);
}

void build_pi_rc_HH(Grid3& g) { g.sym(1, 1, 1, 0.124916); }  // cubici.cxx:468-481

void build_Tij(Grid3& g) {  // cubici.cxx:487-506
  g.sym(2, 2, 1, -0.035140);
  g.sym(2, 2, 2, -0.004048); g.up(2, 2, 2);
}
}  // namespace

// =================================================================== REBO
REBO::REBO() {  // REBO.cxx:817-941
  handled = {H_EL, C_EL};
  Q_[C][C] = 0.313460 * Ao; Q_[C][H] = 0.340776 * Ao; Q_[H][C] = Q_[C][H]; Q_[H][H] = 0.370471 * Ao;
  alpha_[C][C] = 4.7465391 / Ao; alpha_[C][H] = 4.1025498 / Ao; alpha_[H][C] = alpha_[C][H];
  alpha_[H][H] = 3.5362986 / Ao;
  A_[C][C] = 10953.544 * eV; A_[C][H] = 149.94099 * eV; A_[H][C] = A_[C][H]; A_[H][H] = 32.817356 * eV;
  B_[0][C][C] = 12388.792 * eV; B_[0][C][H] = 32.355187 * eV; B_[0][H][C] = B_[0][C][H];
  B_[0][H][H] = 29.632593 * eV;
  B_[1][C][C] = 17.567065 * eV; B_[1][C][H] = 0.0; B_[1][H][C] = 0.0; B_[1][H][H] = 0.0;
  B_[2][C][C] = 30.714932 * eV; B_[2][C][H] = 0.0; B_[2][H][C] = 0.0; B_[2][H][H] = 0.0;
  beta_[0][C][C] = 4.7204523 / Ao; beta_[0][C][H] = 1.4344581 / Ao; beta_[0][H][C] = beta_[0][C][H];
  beta_[0][H][H] = 1.7158922 / Ao;
  beta_[1][C][C] = 1.4332132 / Ao; beta_[1][C][H] = 0.0; beta_[1][H][C] = 0.0; beta_[1][H][H] = 0.0;
  beta_[2][C][C] = 1.3826913 / Ao; beta_[2][C][H] = 0.0; beta_[2][H][C] = 0.0; beta_[2][H][H] = 0.0;
  rho_[C][C] = 0.0; rho_[C][H] = 1.09 * Ao; rho_[H][C] = rho_[C][H]; rho_[H][H] = 0.7415887 * Ao;
  R_[C][C][0] = 1.7 * Ao; R_[H][H][0] = 1.1 * Ao; R_[C][H][0] = 1.3 * Ao; R_[H][C][0] = R_[C][H][0];
  R_[C][C][1] = 2.0 * Ao; R_[H][H][1] = 1.7 * Ao; R_[C][H][1] = 1.8 * Ao; R_[H][C][1] = R_[C][H][1];
  Rp_[C][C][0] = 1.7 * Ao; Rp_[H][H][0] = 1.1 * Ao; Rp_[C][H][0] = 1.3 * Ao; Rp_[H][C][0] = Rp_[C][H][0];
  Rp_[C][C][1] = 2.0 * Ao; Rp_[H][H][1] = 1.7 * Ao; Rp_[C][H][1] = 1.6 * Ao; Rp_[H][C][1] = Rp_[C][H][1];
  Ntot_[0] = 2;
  Ntot_[1] = 3;

  {  // G splines, cubici.cxx:36-95
    const double x1[5] = {-1.0, -2.0 / 3.0, -1.0 / 2.0, -1.0 / 3.0, 1.0};
    const double y1[5] = {-0.010000, 0.028207, 0.052804, 0.097321, 1.000000};
    const double d1[5] = {0.104000, 0.131443, 0.170000, 0.400000, 2.834570};
    const double s1[5] = {0.000000, 0.140229, 0.370000, 1.980000, 10.264700};
    gC1.init(x1, y1, d1, s1, 5);
    const double y2[5] = {-0.010000, 0.028207, 0.052804, 0.097321, 8.000000};
    const double d2[5] = {0.104000, 0.131443, 0.170000, 0.400000, 20.243600};
    const double s2[5] = {0.000000, 0.140229, 0.370000, 1.980000, 43.9336000};
    gC2.init(x1, y2, d2, s2, 5);
    const double xh[4] = {-1.0, -5.0 / 6.0, -1.0 / 2.0, 1.0};
    const double yh[4] = {11.235700, 12.595300, 16.811100, 19.991800};
    const double dh[4] = {0.000000, 13.854300, 8.641230, 0.333013};
    const double sh[4] = {115.115000, 32.361800, -25.061700, -0.474189};
    gH.init(xh, yh, dh, sh, 4);
  }
  {  // P splines, cubici.cxx:112-290
    double z[5][5] = {}, zx[5][5] = {}, zy[5][5] = {};
    z[2][0] = -0.000500; z[3][0] = 0.016125; z[1][1] = -0.010960;
    z[2][1] = 0.006326; z[0][2] = -0.027603; z[1][2] = 0.003180;
    pCC.init(z, zx, zy, true);
    double w[5][5] = {};
    w[1][0] = 0.209337; w[2][0] = -0.064450; w[3][0] = -0.303928; w[0][1] = 0.010000;
    w[1][1] = -0.125123; w[2][1] = -0.298905; w[0][2] = -0.122042; w[1][2] = -0.300529;
    w[0][3] = -0.307585;
    pCH.init(w, zx, zy, true);
  }
  {
    Grid3 a(5, 5, 16), b(5, 5, 16), c(5, 5, 16), t(5, 5, 16);
    build_pi_rc_CC(a);
    build_pi_rc_CH(b);
    build_pi_rc_HH(c);
    build_Tij(t);
    rcCC.init(5, 5, 16, 15, 15, true, a.v, a.vi, a.vj, a.vk);
    rcCH.init(5, 5, 16, 15, 15, true, b.v, b.vi, b.vj, b.vk);
    rcHH.init(5, 5, 16, 15, 15, true, c.v, c.vi, c.vj, c.vk);
    tij.init(5, 5, 16, 15, 15, true, t.v, t.vi, t.vj, t.vk);
  }
  nl.Rcutoff = getRcutoff();
}

double REBO::compute(AtomsArray& atoms) {  // REBO.cxx:32-65
  gl = &atoms;
  countNeighbours(atoms);
  double Ei = 0;
  for (size_t ii = 0; ii < atoms.size(); ii++) {
    Atom& ai = atoms[ii];
    if (!isHandled(ai)) continue;
    auto& nli = NL(ai);
    for (size_t jj = 0; jj < nli.size(); jj++) {
      Atom& aj = atoms[nli[jj]];
      if (ai.globalIndex > aj.globalIndex) continue;
      if (&ai == &aj) continue;
      if (!probablyAreNeighbours(ai, aj)) continue;
      Pair ij(ai, aj, R(0, ai, aj), R(1, ai, aj));
      double VAvar = VA(ij);
      Ei += VR(ij, 1.0);
      if (VAvar != 0.0) {
        double BaverVal = Baver(ij, VAvar);
        Ei += BaverVal * VA(ij, BaverVal);
      }
    }
  }
  return Ei;
}

double REBO::Sprime(double a, double mn, double mx) const {  // REBO.cxx:67-86
  ORC_REQUIRE(mn < mx);
  if (a < mn) return 1.0;
  if (a > mx) return 0.0;
  return (1.0 + std::cos(M_PI * (a - mn) / (mx - mn))) / 2.0;
}

double REBO::dSprime(double a, double mn, double mx) const {  // REBO.cxx:88-106
  ORC_REQUIRE(mn < mx);
  if (a < mn || a > mx) return 0.0;
  return (-(M_PI / (mx - mn)) * std::sin(M_PI * (a - mn) / (mx - mn))) / 2.0;
}

double REBO::fprime(Pair& ij, double V) {  // REBO.cxx:108-124
  double R1 = Rp_[e2i(*ij.a1)][e2i(*ij.a2)][0];
  double R2 = Rp_[e2i(*ij.a1)][e2i(*ij.a2)][1];
  double r = ij.r();
  double val = Sprime(r, R1, R2);
  if (V != 0.0 && r >= R1 && r <= R2) ij.r(dSprime(r, R1, R2) * V);
  return val;
}

double REBO::VR(Pair& ij, double V) {  // REBO.cxx:132-160
  double f = ij.f();
  if (f == 0.0) return 0.0;
  double r = ij.r();
  int a = e2i(*ij.a1), b = e2i(*ij.a2);
  double Q = Q_[a][b], A = A_[a][b], al = alpha_[a][b];
  double t1 = A * std::exp(-al * r);
  double val = (1 + Q / r) * t1;
  if (V != 0) {
    double der = -t1 * (Q + al * r * r + al * r * Q) / (r * r);
    ij.r(der * f * V);
    ij.f(val * V);
  }
  return f * val;
}

double REBO::VA(Pair& ij, double V) {  // REBO.cxx:162-198
  double f = ij.f();
  if (f == 0.0) return 0.0;
  double r = ij.r();
  int a = e2i(*ij.a1), b = e2i(*ij.a2);
  double t1 = B_[0][a][b] * std::exp(-beta_[0][a][b] * r);
  double t2 = B_[1][a][b] * std::exp(-beta_[1][a][b] * r);
  double t3 = B_[2][a][b] * std::exp(-beta_[2][a][b] * r);
  double val = t1 + t2 + t3;
  if (V != 0) {
    double der = (-beta_[0][a][b]) * t1 + (-beta_[1][a][b]) * t2 + (-beta_[2][a][b]) * t3;
    ij.r(-der * f * V);
    ij.f(-val * V);
  }
  return -f * val;
}

double REBO::Baver(Pair& ij, double V) {  // REBO.cxx:200-211
  Pair ji = ij.neg();
  double p1 = p_sigma_pi(ij, V / 2.0);
  double p2 = p_sigma_pi(ji, V / 2.0);
  double rc = pi_rc(ij, V);
  double dh = pi_dh(ij, V);
  return (p1 + p2) / 2.0 + rc + dh;
}

double REBO::ExpTerm(Pair& ij, Pair& ik, double V) {  // REBO.cxx:213-233
  if (ij.a1->ID != H_EL) return 1.0;
  double r_ij = ij.r_alter();
  double r_ik = ik.r();
  double rho_ik = rho_[e2i(*ik.a1)][e2i(*ik.a2)];
  double rho_ij = rho_[e2i(*ij.a1)][e2i(*ij.a2)];
  if (rho_ik == 0.0 || rho_ij == 0.0) throw OracleError("REBO rho. CC");
  double val = std::exp((4.0 / Ao) * ((rho_ik - r_ik) - (rho_ij - r_ij)));
  double dt = val * (4.0 / Ao);
  if (V != 0.0) {
    ij.r_alter(dt * V);
    ik.r(-dt * V);
  }
  return val;
}

double REBO::D(Pair& ij, double V) {  // REBO.cxx:235-262
  double Dij = 1.0;
  Atom& ai = *ij.a1;
  auto& nli = NL(ai);
  for (size_t k = 0; k < nli.size(); k++) {
    Atom& ak = (*gl)[nli[k]];
    if (&ak == ij.a2) continue;
    if (!probablyAreNeighbours(ai, ak)) continue;
    Pair ik(ai, ak, R(0, ai, ak), R(1, ai, ak));
    double E = ExpTerm(ij, ik);
    if (E != 0.0) {
      double Gv = G(ij, ik);
      double fv = ik.f();
      if (V != 0.0) {
        G(ij, ik, fv * E * V);
        ik.f(Gv * E * V);
        ExpTerm(ij, ik, Gv * fv * V);
      }
      Dij += Gv * fv * E;
    }
  }
  Dij += P(ij, V);
  return Dij;
}

double REBO::p_sigma_pi(Pair& ij, double V) {  // REBO.cxx:264-275
  double Dv = D(ij);
  if (V != 0) {
    double der = -0.5 * std::pow(Dv, -0.5 - 1.0);
    D(ij, der * V);
  }
  return std::pow(Dv, -0.5);
}

double REBO::G(Pair& ij, Pair& ik, double V) {  // REBO.cxx:277-326
  double c = CosTheta(ij, ik);
  if (ij.a1->ID == C_EL) {
    double N = Nt(ij);
    double Sp = Sprime(N, Ntot_[0], Ntot_[1]);
    double g1 = gC1(c), g2 = gC2(c);
    if (V != 0.0) {
      double dSp = dSprime(N, Ntot_[0], Ntot_[1]);
      double g1d = gC1.d(c), g2d = gC2.d(c);
      if (g2d != 0.0 || g1d != 0.0) CosTheta(ij, ik, g2d * V + g1d * Sp * V + (-g2d * Sp * V));
      if ((g2 != 0.0 || g1 != 0.0) && dSp != 0.0) dfThem(Nts2t, ij, g1 * dSp * V + (-g2 * dSp * V));
    }
    return g2 + Sp * (g1 - g2);
  } else if (ij.a1->ID == H_EL) {
    if (V != 0.0) {
      double gd = gH.d(c);
      if (gd != 0.0) CosTheta(ij, ik, gd * V);
    }
    return gH(c);
  }
  throw OracleError("REBO::G() : unknown element");
}

// REBO.cxx:328-393.  Per-atom coordination sums over the REBO neighbour list
// and the lists of neighbours inside the switching zone ("touch" lists).
void REBO::countNeighbours(AtomsArray& atoms) {
  size_t n = atoms.size();
  Nts.assign(n, 0.0);
  NCs.assign(n, 0.0);
  NHs.assign(n, 0.0);
  Nts2t.assign(n, {});
  NCs2t.assign(n, {});
  NHs2t.assign(n, {});
  for (size_t i = 0; i < n; i++) {
    Atom& ai = atoms[i];
    if (!isHandled(ai)) continue;
    auto& nli = NL(ai);
    for (size_t k = 0; k < nli.size(); k++) {
      Atom& ak = atoms[nli[k]];
      double fval = 0.0;
      bool touch = false;
      double r = norm(depos(ai, ak));
      double R2 = R(1, ai, ak);
      if (r > R2) continue;
      double R1 = R(0, ai, ak);
      if (r < R1) {
        fval = 1.0;
      } else {
        touch = true;
        fval = (1.0 + std::cos(M_PI * (r - R1) / (R2 - R1))) / 2.0;
      }
      Nts[i] += fval;
      if (touch) Nts2t[i].push_back(nli[k]);
      if (ak.ID == H_EL) {
        NHs[i] += fval;
        if (touch) NHs2t[i].push_back(nli[k]);
      }
      if (ak.ID == C_EL) {
        NCs[i] += fval;
        if (touch) NCs2t[i].push_back(nli[k]);
      }
    }
  }
}

void REBO::dfThem(const std::vector<std::vector<int>>& pa, Pair& ij, double V) {  // REBO.cxx:408-421
  if (V == 0.0) return;
  const auto& they = pa[ij.a1->globalIndex];
  for (size_t k = 0; k < they.size(); k++) {
    Atom& ak = (*gl)[they[k]];
    if (&ak == ij.a2) continue;
    Pair ik(*ij.a1, ak, R(0, *ij.a1, ak), R(1, *ij.a1, ak));
    ik.f(V);
  }
}

double REBO::Nt(Pair& ij, double V) {  // REBO.cxx:423-438
  if (V != 0.0) dfThem(Nts2t, ij, V);
  double N = Nts[ij.a1->globalIndex];
  if (ij.r() < R(1, *ij.a1, *ij.a2)) N -= ij.f();
  ORC_REQUIRE(N > -1e-15);
  if (N < 0) N = 0;
  return N;
}

double REBO::NH(Pair& ij, double V) {  // REBO.cxx:450-466
  if (V != 0.0) dfThem(NHs2t, ij, V);
  double N = NHs[ij.a1->globalIndex];
  if (ij.a2->ID == H_EL && ij.r() < R(1, *ij.a1, *ij.a2)) N -= ij.f();
  ORC_REQUIRE(N > -1e-15);
  if (N < 0) N = 0;
  return N;
}

double REBO::NC(Pair& ij, double V) {  // REBO.cxx:477-493
  if (V != 0.0) dfThem(NCs2t, ij, V);
  double N = NCs[ij.a1->globalIndex];
  if (ij.a2->ID == C_EL && ij.r() < R(1, *ij.a1, *ij.a2)) N -= ij.f();
  ORC_REQUIRE(N > -1e-15);
  if (N < 0) N = 0;
  return N;
}

// NconjSum1 / NconjSum2, REBO.cxx:540-608: sum over carbon neighbours k of i
// (or l of j) of f_ki * S'(N_ki; 2, 3).
double REBO::NconjSum(Pair& ij, bool second, double V) {
  double sum = 0.0;
  Atom& c = second ? *ij.a2 : *ij.a1;
  Atom& other = second ? *ij.a1 : *ij.a2;
  auto& nlc = NL(c);
  for (size_t k = 0; k < nlc.size(); k++) {
    Atom& ak = (*gl)[nlc[k]];
    if (&ak == &other || ak.ID != C_EL) continue;
    if (!probablyAreNeighbours(c, ak)) continue;
    Pair ki(ak, c, R(0, c, ak), R(1, c, ak));
    double f_ki = ki.f();
    double N = Nt(ki);
    double Sp = Sprime(N, Ntot_[0], Ntot_[1]);
    if (V != 0.0) {
      ki.f(Sp * V);
      dfThem(Nts2t, ki, f_ki * dSprime(N, Ntot_[0], Ntot_[1]) * V);
    }
    sum += f_ki * Sp;
  }
  return sum;
}

double REBO::Nconj(Pair& ij, double V) {  // REBO.cxx:610-624
  double N = 1.0;
  double s1 = NconjSum(ij, false);
  if (V != 0.0 && s1 != 0.0) NconjSum(ij, false, 2.0 * s1 * V);
  N += SQR(s1);
  double s2 = NconjSum(ij, true);
  if (V != 0.0 && s2 != 0.0) NconjSum(ij, true, 2.0 * s2 * V);
  N += SQR(s2);
  return N;
}

double REBO::P(Pair& ij, double V) {  // REBO.cxx:626-657
  Func2O* f = nullptr;
  if (ij.a1->ID == C_EL && ij.a2->ID == C_EL) f = &pCC;
  else if (ij.a1->ID == C_EL && ij.a2->ID == H_EL) f = &pCH;
  else return 0.0;
  double nh = NH(ij), nc = NC(ij);
  double val = f->eval(nh, nc, 0);
  if (V != 0.0) {
    dfThem(NHs2t, ij, f->eval(nh, nc, 1) * V);
    dfThem(NCs2t, ij, f->eval(nh, nc, 2) * V);
  }
  return val;
}

double REBO::pi_rc(Pair& ij, double V) {  // REBO.cxx:659-714
  Pair ji = ij.neg();
  Func3O* f = nullptr;
  if (ij.a1->ID == C_EL && ij.a2->ID == C_EL) f = &rcCC;
  else if (ij.a1->ID == C_EL && ij.a2->ID == H_EL) f = &rcCH;
  else if (ij.a1->ID == H_EL && ij.a2->ID == C_EL) return pi_rc(ji, V);
  else if (ij.a1->ID == H_EL && ij.a2->ID == H_EL) f = &rcHH;
  else return 0.0;
  double Ni = Nt(ij), Nj = Nt(ji), Nc = Nconj(ij);
  double val = f->eval(Ni, Nj, Nc, 0);
  if (V != 0.0) {
    dfThem(Nts2t, ij, f->eval(Ni, Nj, Nc, 1) * V);
    dfThem(Nts2t, ji, f->eval(Ni, Nj, Nc, 2) * V);
    Nconj(ij, f->eval(Ni, Nj, Nc, 3) * V);
  }
  return val;
}

double REBO::pi_dh_sum(Pair& ij, double V) {  // REBO.cxx:716-765
  double s = 0.0;
  Atom& ai = *ij.a1;
  Atom& aj = *ij.a2;
  auto& nli = NL(ai);
  auto& nlj = NL(aj);
  for (size_t k = 0; k < nli.size(); k++) {
    Atom& ak = (*gl)[nli[k]];
    if (&ak == &aj) continue;
    if (!probablyAreNeighbours(ai, ak)) continue;
    for (size_t l = 0; l < nlj.size(); l++) {
      Atom& al = (*gl)[nlj[l]];
      if (&al == &ai) continue;
      if (!probablyAreNeighbours(aj, al)) continue;
      if (&ak == &al) continue;
      Pair ik(ak, ai, R(0, ai, ak), R(1, ai, ak));
      if (std::fabs(SinTheta(ij, ik)) < 0.1) continue;
      Pair ki = ik.neg();
      double f_ik = fprime(ik);
      Pair jl(al, aj, R(0, aj, al), R(1, aj, al));
      if (std::fabs(SinTheta(ij, jl)) < 0.1) continue;
      Pair lj = jl.neg();
      double f_jl = fprime(jl);
      double c = -CosDihedral(ij, ki, lj);
      if (V != 0) {
        CosDihedral(ij, ki, lj, +1.0 * 2 * c * f_ik * f_jl * V);
        fprime(ik, (1.0 - SQR(c)) * f_jl * V);
        fprime(jl, (1.0 - SQR(c)) * f_ik * V);
      }
      s += (1.0 - SQR(c)) * f_ik * f_jl;
    }
  }
  return s;
}

double REBO::pi_dh_Tij(Pair& ij, double V) {  // REBO.cxx:767-800
  Pair ji = ij.neg();
  if (ij.a1->ID == H_EL || ij.a2->ID == H_EL) return 0.0;
  double Ni = Nt(ij), Nj = Nt(ji), Nc = Nconj(ij);
  double val = tij.eval(Ni, Nj, Nc, 0);
  if (V != 0.0) {
    double di = tij.eval(Ni, Nj, Nc, 1);
    double dj = tij.eval(Ni, Nj, Nc, 2);
    double dk = tij.eval(Ni, Nj, Nc, 3);
    if (val == 0.0 && di == 0.0 && dj == 0.0 && dk == 0.0) return 0.0;
    if (di != 0.0) dfThem(Nts2t, ij, di * V);
    if (dj != 0.0) dfThem(Nts2t, ji, dj * V);
    if (dk != 0.0) Nconj(ij, dk * V);
  }
  return val;
}

double REBO::pi_dh(Pair& ij, double V) {  // REBO.cxx:802-814
  double T = pi_dh_Tij(ij);
  if (T == 0.0) return 0.0;
  double s = pi_dh_sum(ij);
  if (V != 0.0) {
    if (s != 0.0) pi_dh_Tij(ij, s * V);
    pi_dh_sum(ij, T * V);
  }
  return T * s;
}

// =================================================================== AIREBO LJ
AIREBO_LJ::AIREBO_LJ(REBO& r) : rebo(r) {  // AIREBO_LJ.cxx:219-287
  handled = {H_EL, C_EL};
  const int C = 0, H = 1;
  sigma_[C][C] = 3.40 * Ao;
  sigma_[H][H] = 2.65 * Ao;
  sigma_[C][H] = 0.5 * (sigma_[C][C] + sigma_[H][H]);
  sigma_[H][C] = sigma_[C][H];
  zeta_[C][C] = 0.00284 * eV;
  zeta_[H][H] = 0.00150 * eV;
  zeta_[C][H] = std::sqrt(zeta_[C][C] * zeta_[H][H]);
  zeta_[H][C] = zeta_[C][H];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      RLJ_[a][b][0] = sigma_[a][b];
      RLJ_[a][b][1] = sigma_[a][b] * std::pow(2.0, 1.0 / 6.0);
    }
  R_[C][C][0] = 5.0 * Ao; R_[H][H][0] = 4.0 * Ao; R_[C][H][0] = 4.5 * Ao; R_[H][C][0] = R_[C][H][0];
  R_[C][C][1] = R_[C][C][0] + 0.5 * Ao;
  R_[H][H][1] = R_[H][H][0] + 0.5 * Ao;
  R_[C][H][1] = R_[C][H][0] + 0.5 * Ao;
  R_[H][C][1] = R_[C][H][1];
  b_[C][C][0] = 0.77; b_[H][H][0] = 0.32; b_[C][H][0] = 0.75; b_[H][C][0] = b_[C][H][0];
  b_[C][C][1] = 0.81; b_[H][H][1] = 0.42; b_[C][H][1] = 0.90; b_[H][C][1] = b_[C][H][1];
  nl.Rcutoff = getRcutoff();
}

double AIREBO_LJ::compute(AtomsArray& atoms) {  // AIREBO_LJ.cxx:34-77
  fill_Cij(atoms);
  double Ei = 0;
  for (size_t ii = 0; ii < atoms.size(); ii++) {
    Atom& ai = atoms[ii];
    if (!isHandled(ai)) continue;
    auto& nli = NL(ai);
    for (size_t jj = 0; jj < nli.size(); jj++) {
      Atom& aj = atoms[nli[jj]];
      if (ai.globalIndex > aj.globalIndex) continue;
      if (&ai == &aj) continue;
      if (!probablyAreNeighbours(ai, aj)) continue;
      Pair ij(ai, aj, R(0, ai, aj), R(1, ai, aj));
      const double V = 1.0;
      double C = Cij(ij);
      if (C > 0.0) {
        double SrSb = StrStb(ij);
        double LJ = VLJ(ij, SrSb * C * V);
        StrStb(ij, C * LJ * V);
        Cij(ij, SrSb * LJ * V);
        Ei += SrSb * C * LJ;
      }
    }
  }
  return Ei;
}

double AIREBO_LJ::S(double a, double mn, double mx) const {  // AIREBO_LJ.cxx:116-135
  ORC_REQUIRE(mn < mx);
  if (a < mn) return 1.0;
  if (a > mx) return 0.0;
  double t = (a - mn) / (mx - mn);
  return 1.0 - t * t * (3.0 - 2.0 * t);
}

double AIREBO_LJ::dS(double a, double mn, double mx) const {  // AIREBO_LJ.cxx:137-156
  ORC_REQUIRE(mn < mx);
  if (a < mn || a > mx) return 0.0;
  double t = (a - mn) / (mx - mn);
  return (-2.0 * t * (3.0 - 2.0 * t) + 2.0 * t * t) / (mx - mn);
}

double AIREBO_LJ::StrStb(Pair& ij, double V) {  // AIREBO_LJ.cxx:79-114
  double val = 1.0;
  int a = rebo.e2i(*ij.a1), b = rebo.e2i(*ij.a2);
  double r = ij.r();
  double rmin = RLJ_[a][b][0], rmax = RLJ_[a][b][1];
  double Str = S(r, rmin, rmax);
  if (Str != 0.0) {
    double bij = BijAsterix(ij);
    double bmin = b_[a][b][0], bmax = b_[a][b][1];
    double Stb = S(bij, bmin, bmax);
    if (V != 0.0) {
      double dStr = dS(r, rmin, rmax);
      if (dStr != 0.0) ij.r(-dStr * V);
      double dStb = dS(bij, bmin, bmax);
      if (dStr != 0.0 && Stb != 0.0) ij.r(dStr * Stb * V);
      if (dStb != 0.0) BijAsterix(ij, Str * dStb * V);
    }
    val += Str * Stb - Str;
  }
  return val;
}

double AIREBO_LJ::VLJ(Pair& ij, double V) {  // AIREBO_LJ.cxx:158-189
  double f = ij.f();
  if (f == 0.0) return 0.0;
  double r = ij.r();
  int a = rebo.e2i(*ij.a1), b = rebo.e2i(*ij.a2);
  double s = sigma_[a][b] / r;
  double s6 = s;
  for (int i = 2; i <= 6; i++) s6 *= s;
  double s12 = s6 * s6;
  double z = zeta_[a][b];
  double val = 4.0 * z * (s12 - s6);
  if (V != 0.0) {
    double der = 4.0 * z * (-12.0 * s12 / r + 6.0 * s6 / r);
    ij.r(der * f * V);
    ij.f(val * V);
  }
  return f * val;
}

// b*_ij: the REBO bond order at the actual geometry of (i,j) with
// r_ij := R0 inside ExpTerm only (AIREBO_LJ.cxx:191-202).
double AIREBO_LJ::BijAsterix(Pair& ij, double V) {
  double R0 = rebo.R(0, *ij.a1, *ij.a2);
  Pair s(*ij.a1, *ij.a2, R0, rebo.R(1, *ij.a1, *ij.a2), R0);
  return rebo.Baver(s, V);
}

// AIREBO_LJ.cxx:289-358: best path weight (product of REBO cutoff values)
// over paths of at most three bonds, with the reference's tie rules.
void AIREBO_LJ::fill_Cij(AtomsArray& atoms) {
  size_t n = atoms.size();
  CA.assign(n, {});
  REBO& R = rebo;
  for (size_t i1 = 0; i1 < n; i1++) {
    Atom& a1 = atoms[i1];
    if (!isHandled(a1)) continue;
    auto& nl1 = R.NL(a1);
    for (size_t i2 = 0; i2 < nl1.size(); i2++) {
      Atom& a2 = atoms[nl1[i2]];
      if (!R.probablyAreNeighbours(a1, a2)) continue;
      Pair p1(a1, a2, R.R(0, a1, a2), R.R(1, a1, a2));
      {
        CEl& pa = CA[a1.globalIndex][a2.globalIndex];
        double w = pa.w;
        double nw = p1.f();
        if (nw > w || (w > 0.0 && nw == w && pa.path.size() > 1)) {
          pa.path.clear();
          pa.w = nw;
          pa.path.push_back(p1);
        }
      }
      auto& nl2 = R.NL(a2);
      for (size_t i3 = 0; i3 < nl2.size(); i3++) {
        Atom& a3 = atoms[nl2[i3]];
        if (!R.probablyAreNeighbours(a2, a3)) continue;
        Pair p2(a2, a3, R.R(0, a2, a3), R.R(1, a2, a3));
        {
          CEl& pa = CA[a1.globalIndex][a3.globalIndex];
          double w = pa.w;
          double nw = p1.f() * p2.f();
          if (nw > w || (w > 0.0 && nw == w && pa.path.size() > 2)) {
            pa.path.clear();
            pa.w = nw;
            pa.path.push_back(p1);
            pa.path.push_back(p2);
          }
        }
        auto& nl3 = R.NL(a3);
        for (size_t i4 = 0; i4 < nl3.size(); i4++) {
          Atom& a4 = atoms[nl3[i4]];
          if (!R.probablyAreNeighbours(a3, a4)) continue;
          Pair p3(a3, a4, R.R(0, a3, a4), R.R(1, a3, a4));
          CEl& pa = CA[a1.globalIndex][a4.globalIndex];
          double w = pa.w;
          double nw = p1.f() * p2.f() * p3.f();
          if (nw > w) {
            pa.path.clear();
            pa.w = nw;
            pa.path.push_back(p1);
            pa.path.push_back(p2);
            pa.path.push_back(p3);
          }
        }
      }
    }
  }
}

double AIREBO_LJ::Cij(Pair& ij, double V) {  // AIREBO_LJ.cxx:382-408
  auto& m = CA[ij.a1->globalIndex];
  auto it = m.find(ij.a2->globalIndex);
  double w = 0.0;
  if (it != m.end()) {
    w = it->second.w;
    if (V != 0.0) {
      auto& path = it->second.path;
      for (size_t p = 0; p < path.size(); ++p) {
        double mm = 1.0;
        for (size_t q = 0; q < path.size(); ++q)
          if (p != q) mm *= path[q].f();
        path[p].f(-mm * V);
      }
    }
  }
  return 1 - w;
}

// =================================================================== ETors
ETors::ETors() {  // AIREBO_ETors.cxx:117-133
  zetaCC_[C][C] = 0.3079 * eV;
  zetaCC_[H][H] = 0.1250 * eV;
  zetaCC_[C][H] = 0.1787 * eV;
  zetaCC_[H][C] = zetaCC_[C][H];
  nl.Rcutoff = getRcutoff();
}

double ETors::compute(AtomsArray& atoms) {  // AIREBO_ETors.cxx:32-98
  gl = &atoms;
  double Ei = 0;
  for (size_t ii = 0; ii < atoms.size(); ii++) {
    Atom& ai = atoms[ii];
    if (!isHandled(ai)) continue;
    auto& nli = NL(ai);
    for (size_t jj = 0; jj < nli.size(); jj++) {
      Atom& aj = atoms[nli[jj]];
      if (ai.globalIndex > aj.globalIndex) continue;
      if (&ai == &aj) continue;
      if (!probablyAreNeighbours(ai, aj)) continue;
      Pair ij(ai, aj, R(0, ai, aj), R(1, ai, aj));
      for (size_t k = 0; k < nli.size(); k++) {
        Atom& ak = atoms[nli[k]];
        if (&ak == &ai || &ak == &aj) continue;
        if (!probablyAreNeighbours(ai, ak)) continue;
        Pair ki(ak, ai, R(0, ak, ai), R(1, ak, ai));
        auto& nlj = NL(aj);
        for (size_t l = 0; l < nlj.size(); l++) {
          Atom& al = atoms[nlj[l]];
          if (&al == &ai || &al == &aj || &al == &ak) continue;
          if (!probablyAreNeighbours(aj, al)) continue;
          Pair jl(aj, al, R(0, aj, al), R(1, aj, al));
          if (std::fabs(SinTheta(ij, ki)) < 0.1) continue;
          if (std::fabs(SinTheta(ij, jl)) < 0.1) continue;
          const double V = 1.0;
          Pair ik = ki.neg();
          double wki = ki.f(), wij = ij.f(), wjl = jl.f();
          double vt = Vtors(ij, ik, jl, wki * wij * wjl * V);
          ki.f(wij * wjl * vt * V);
          ij.f(wki * wjl * vt * V);
          jl.f(wki * wij * vt * V);
          Ei += wki * wij * wjl * vt;
        }
      }
    }
  }
  return Ei;
}

double ETors::Vtors(Pair& ij, Pair& ik, Pair& jl, double V) {  // AIREBO_ETors.cxx:100-115
  if (ij.a1->ID != C_EL || ij.a2->ID != C_EL) return 0.0;
  double c = -CosDihedral(ij, ik, jl);
  double z = zetaCC_[e2i(*ik.a2)][e2i(*jl.a2)];
  double x = 0.5 * (1.0 + c);
  double val = (256.0 / 405.0) * z * std::pow(x, 5.0) - 0.1 * z;
  if (val != 0) {
    double der = (256.0 / 405.0) * z * std::pow(x, 5.0 - 1.0) * 0.5;
    if (!compat) der *= 5.0;  // the true derivative of x^5
    CosDihedral(ij, ik, jl, -1.0 * der * V);
  }
  return val;
}

// =================================================================== Brenner
Brenner::Brenner(int ps) : paramSet(ps) {  // Brenner.cxx:453-611
  handled = {H_EL, C_EL};
  if (ps == 0) {
    Re_[C][C] = 1.315 * Ao; Re_[H][H] = 0.74144 * Ao; Re_[C][H] = 1.1199 * Ao;
    De_[C][C] = 6.325 * eV; De_[H][H] = 4.7509 * eV; De_[C][H] = 3.6422 * eV;
    beta_[C][C] = 1.5 / Ao; beta_[H][H] = 1.9436 / Ao; beta_[C][H] = 1.9583 / Ao;
    S_[C][C] = 1.29; S_[H][H] = 2.3432; S_[C][H] = 1.7386;
    delta_[C][C] = 0.80469; delta_[H][H] = 0.80469; delta_[C][H] = 0.0;
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++)
        for (int z = 0; z < 2; z++) alpha_[x][y][z] = 3.0 / Ao;
    ao_ = 0.011304; co_ = 19.0; do_ = 2.5; G_HH_ = 4.0;
  } else {
    Re_[C][C] = 1.39 * Ao; Re_[H][H] = 0.74144 * Ao; Re_[C][H] = 1.1199 * Ao;
    De_[C][C] = 6.0 * eV; De_[H][H] = 4.7509 * eV; De_[C][H] = 3.6422 * eV;
    beta_[C][C] = 2.1 / Ao; beta_[H][H] = 1.9436 / Ao; beta_[C][H] = 1.9583 / Ao;
    S_[C][C] = 1.22; S_[H][H] = 2.3432; S_[C][H] = 1.69077;
    delta_[C][C] = 0.5; delta_[H][H] = 0.5; delta_[C][H] = 0.0;
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++)
        for (int z = 0; z < 2; z++) alpha_[x][y][z] = 4.0 / Ao;
    ao_ = 0.00020813; co_ = 330.0; do_ = 3.5; G_HH_ = 12.33;
  }
  alpha_[C][H][C] = 0.0; alpha_[C][C][H] = 0.0; alpha_[C][C][C] = 0.0;
  Re_[H][C] = Re_[C][H]; De_[H][C] = De_[C][H]; beta_[H][C] = beta_[C][H]; S_[H][C] = S_[C][H];
  delta_[H][C] = delta_[C][H];
  R_[C][C][0] = 1.7 * Ao; R_[H][H][0] = 1.1 * Ao; R_[C][H][0] = 1.3 * Ao; R_[H][C][0] = R_[C][H][0];
  R_[C][C][1] = 2.0 * Ao; R_[H][H][1] = 1.7 * Ao; R_[C][H][1] = 1.8 * Ao; R_[H][C][1] = R_[C][H][1];

  // Brenner/cubici.cxx:36-346
  double hcc[5][5] = {}, hch[5][5] = {}, dhh[5][5] = {}, dhc[5][5] = {}, zero[5][5] = {};
  if (ps == 0) {
    hcc[1][1] = -0.0175; hcc[2][0] = -0.0070; hcc[3][0] = 0.0119; hcc[1][2] = 0.0115; hcc[2][1] = 0.0118;
    hch[1][0] = -0.0760; hch[2][0] = -0.2163; hch[3][0] = -0.3375; hch[0][1] = -0.1792;
    hch[0][2] = -0.2407; hch[1][1] = -0.2477; hch[2][1] = -0.3320; hch[0][3] = -0.3323;
    hch[1][2] = -0.3321;
    dhh[2][0] = -0.13075; dhh[1][1] = -0.0764; dhc[0][2] = -0.07655; dhc[1][1] = -0.12805;
  } else {
    hcc[1][1] = -0.0226; hcc[2][0] = -0.0061; hcc[3][0] = 0.0173; hcc[1][2] = 0.0149; hcc[2][1] = 0.0160;
    hch[1][0] = -0.0984; hch[2][0] = -0.2878; hch[3][0] = -0.4507; hch[0][1] = -0.2479;
    hch[0][2] = -0.3221; hch[1][1] = -0.3344; hch[2][1] = -0.4438; hch[0][3] = -0.4460;
    hch[1][2] = -0.4449;
    dhh[2][0] = -0.17615; dhh[1][1] = -0.09795; dhc[0][2] = -0.09905; dhc[1][1] = -0.17325;
  }
  hCC.init(hcc, zero, zero, false);
  hCH.init(hch, dhh, dhc, false);
  Grid3 g(5, 5, 6);
  auto F = [&](int i, int j, int k, double v) { g.at(g.v, i, j, k) = v; };
  auto DI = [&](int i, int j, int k, double v) { g.at(g.vi, i, j, k) = v; };
  auto DJ = [&](int i, int j, int k, double v) { g.at(g.vj, i, j, k) = v; };
  if (ps == 0) {
    F(2, 3, 1, -0.0465); F(3, 2, 1, -0.0465); F(2, 3, 2, -0.0465); F(3, 2, 2, -0.0465);
    F(1, 2, 2, -0.0355); F(2, 1, 2, -0.0355); F(1, 2, 1, 0.0126); F(2, 1, 1, 0.0126);
    F(1, 3, 1, -0.1130); F(3, 1, 1, -0.1130); F(1, 3, 2, -0.1130); F(3, 1, 2, -0.1130);
    F(0, 3, 1, -0.1220); F(3, 0, 1, -0.1220); F(0, 3, 2, -0.1220); F(3, 0, 2, -0.1220);
    F(0, 2, 2, -0.0445); F(2, 0, 2, -0.0445); F(0, 2, 1, 0.0320); F(2, 0, 1, 0.0320);
    F(0, 1, 1, 0.1100); F(1, 0, 1, 0.1100);
    F(1, 1, 2, 0.0074); F(1, 1, 1, 0.1511); F(2, 2, 1, 0.0750);
    DI(2, 0, 1, -0.1160); DI(2, 1, 1, -0.13205); DI(2, 0, 2, -0.0610); DI(1, 2, 2, 0.02225);
    DI(1, 3, 2, 0.03775); DI(2, 3, 2, 0.0565); DI(2, 3, 1, 0.0565); DI(2, 1, 2, -0.0602);
    DJ(0, 2, 1, -0.1160); DJ(1, 2, 1, -0.13205); DJ(0, 2, 2, -0.0610); DJ(2, 1, 2, 0.02225);
    DJ(3, 1, 2, 0.03775); DJ(3, 2, 2, 0.0565); DJ(3, 2, 1, 0.0565); DJ(1, 2, 2, -0.0602);
  } else {
    F(2, 3, 1, -0.0363); F(3, 2, 1, -0.0363); F(2, 3, 2, -0.0363); F(3, 2, 2, -0.0363);
    F(1, 2, 2, -0.0243); F(2, 1, 2, -0.0243); F(1, 2, 1, 0.0120); F(2, 1, 1, 0.0120);
    F(1, 3, 1, -0.0903); F(3, 1, 1, -0.0903); F(1, 3, 2, -0.0903); F(3, 1, 2, -0.0903);
    F(0, 3, 1, -0.0904); F(3, 0, 1, -0.0904); F(0, 3, 2, -0.0904); F(3, 0, 2, -0.0904);
    F(0, 2, 2, -0.0269); F(2, 0, 2, -0.0269); F(0, 2, 1, 0.0427); F(2, 0, 1, 0.0427);
    F(0, 1, 1, 0.0996); F(1, 0, 1, 0.0996);
    F(1, 1, 2, 0.0108); F(1, 1, 1, 0.1264); F(2, 2, 1, 0.0605);
    DI(2, 0, 1, -0.0950); DI(2, 1, 1, -0.10835); DI(2, 0, 2, -0.0452); DI(1, 2, 2, 0.01345);
    DI(1, 3, 2, 0.02705); DI(2, 3, 2, 0.04515); DI(2, 3, 1, 0.04515); DI(2, 1, 2, -0.05055);
    DJ(0, 2, 1, -0.0950); DJ(1, 2, 1, -0.10835); DJ(0, 2, 2, -0.0452); DJ(2, 1, 2, 0.01345);
    DJ(3, 1, 2, 0.02705); DJ(3, 2, 2, 0.04515); DJ(3, 2, 1, 0.04515); DJ(1, 2, 2, -0.05055);
  }
  for (int i = 0; i < 5; i++)  // FCC[i][j][k>=3] = FCC[i][j][2], Brenner/cubici.cxx:300-305
    for (int j = 0; j < 5; j++)
      for (int k = 3; k < 6; k++) g.at(g.v, i, j, k) = g.at(g.v, i, j, 2);
  std::vector<double> zk(g.v.size(), 0.0);
  fF.init(5, 5, 6, 5, 4, false, g.v, g.vi, g.vj, zk);
  nl.Rcutoff = getRcutoff();
}

double Brenner::compute(AtomsArray& atoms) {  // Brenner.cxx:31-61
  double Ei = 0;
  gl_ = &atoms;
  for (size_t ii = 0; ii < atoms.size(); ii++) {
    Atom& ai = atoms[ii];
    if (!isHandled(ai)) continue;
    auto& nli = NL(ai);
    for (size_t jj = 0; jj < nli.size(); jj++) {
      Atom& aj = atoms[nli[jj]];
      if (ai.globalIndex > aj.globalIndex) continue;
      if (!isHandled(aj)) continue;
      if (&ai == &aj) continue;
      if (!probablyAreNeighbours(ai, aj)) continue;
      Pair ij(ai, aj, R(0, ai, aj), R(1, ai, aj));
      double VAvar = VA(ij);
      Ei += VR(ij, 1.0);
      if (VAvar != 0.0) {
        double b = Baver(ij, -VAvar);
        Ei += -b * VA(ij, -b);
      }
    }
  }
  return Ei;
}

double Brenner::VR(Pair& ij, double V) {  // Brenner.cxx:63-94
  double f = ij.f();
  if (f == 0.0) return 0.0;
  double r = ij.r();
  int a = e2i(*ij.a1), b = e2i(*ij.a2);
  double A = De_[a][b] / (S_[a][b] - 1.0);
  double B = std::sqrt(2.0 * S_[a][b]) * beta_[a][b];
  double val = A * std::exp(-B * (r - Re_[a][b]));
  if (V != 0) {
    ij.r(val * (-B) * f * V);
    ij.f(val * V);
  }
  return f * val;
}

double Brenner::VA(Pair& ij, double V) {  // Brenner.cxx:96-127
  double f = ij.f();
  if (f == 0.0) return 0.0;
  double r = ij.r();
  int a = e2i(*ij.a1), b = e2i(*ij.a2);
  double A = De_[a][b] * S_[a][b] / (S_[a][b] - 1.0);
  double B = std::sqrt(2.0 / S_[a][b]) * beta_[a][b];
  double val = A * std::exp(-B * (r - Re_[a][b]));
  if (V != 0) {
    ij.r(val * (-B) * f * V);
    ij.f(val * V);
  }
  return f * val;
}

double Brenner::Baver(Pair& ij, double V) {  // Brenner.cxx:129-141
  Pair ji = ij.neg();
  double b1 = B(ij, V / 2.0);
  double b2 = B(ji, V / 2.0);
  double fn = FN(ij, V / 2.0);
  return (b1 + b2) / 2.0 + (fn / 2.0);
}

double Brenner::ExpTerm(Pair& ij, Pair& ik, double V) {  // Brenner.cxx:143-166
  double al = alpha_[e2i(*ij.a1)][e2i(*ij.a2)][e2i(*ik.a2)];
  if (al == 0.0) return 1.0;
  double r_ij = ij.r(), r_ik = ik.r();
  double val = std::exp(al * ((r_ij - Re_[e2i(*ij.a1)][e2i(*ij.a2)]) - (r_ik - Re_[e2i(*ik.a1)][e2i(*ik.a2)])));
  double dt = val * al;
  if (V != 0.0) {
    ij.r(dt * V);
    ik.r(-dt * V);
  }
  return val;
}

double Brenner::D(Pair& ij, double V) {  // Brenner.cxx:168-196
  double Dij = 1.0;
  Atom& ai = *ij.a1;
  auto& nli = NL(ai);
  for (size_t k = 0; k < nli.size(); k++) {
    Atom& ak = (*gl_)[nli[k]];
    if (!isHandled(ak)) continue;
    if (&ak == ij.a2) continue;
    if (!probablyAreNeighbours(ai, ak)) continue;
    Pair ik(ai, ak, R(0, ai, ak), R(1, ai, ak));
    double E = ExpTerm(ij, ik);
    if (E != 0.0) {
      double Gv = G(ij, ik);
      double fv = ik.f();
      if (V != 0.0) {
        G(ij, ik, fv * E * V);
        ik.f(Gv * E * V);
        ExpTerm(ij, ik, Gv * fv * V);
      }
      Dij += Gv * fv * E;
    }
  }
  Dij += HN(ij, V);
  return Dij;
}

double Brenner::B(Pair& ij, double V) {  // Brenner.cxx:198-209
  double Dv = D(ij);
  double dl = delta_[e2i(*ij.a1)][e2i(*ij.a1)];
  if (V != 0) {
    double der = -dl * std::pow(Dv, -dl - 1.0);
    D(ij, der * V);
  }
  return std::pow(Dv, -dl);
}

double Brenner::G(Pair& ij, Pair& ik, double V) {  // Brenner.cxx:211-243
  if (ij.a1->ID == C_EL) {
    double c = CosTheta(ij, ik);
    if (V != 0) {
      double dG = 2.0 * ao_ * SQR(co_) * (1.0 + c) / SQR(SQR(do_) + SQR(1.0 + c));
      CosTheta(ij, ik, dG * V);
    }
    return ao_ * (1.0 + SQR(co_) / SQR(do_) - SQR(co_) / (SQR(do_) + SQR(1.0 + c)));
  } else if (ij.a1->ID == H_EL) {
    return G_HH_;
  }
  throw OracleError("Brenner::G() : unknown element");
}

double Brenner::Nt(Pair& ij, double V) {  // Brenner.cxx:245-262
  ORC_REQUIRE(ij.a1->ID == C_EL);
  double s = 0;
  Atom& ai = *ij.a1;
  auto& nli = NL(ai);
  for (size_t k = 0; k < nli.size(); k++) {
    Atom& ak = (*gl_)[nli[k]];
    if (!isHandled(ak)) continue;
    if (&ak == ij.a2) continue;
    if (!probablyAreNeighbours(ai, ak)) continue;
    Pair ik(ai, ak, R(0, ai, ak), R(1, ai, ak));
    s += ik.f(V);
  }
  return s;
}

double Brenner::NH(Pair& ij, double V) {  // Brenner.cxx:264-280
  ORC_REQUIRE(ij.a1->ID == C_EL);
  double s = 0;
  Atom& ai = *ij.a1;
  auto& nli = NL(ai);
  for (size_t k = 0; k < nli.size(); k++) {
    Atom& ak = (*gl_)[nli[k]];
    if (ak.ID != H_EL || &ak == ij.a2) continue;
    if (!probablyAreNeighbours(ai, ak)) continue;
    Pair ik(ai, ak, R(0, ai, ak), R(1, ai, ak));
    s += ik.f(V);
  }
  return s;
}

double Brenner::NC(Pair& ij, double V) {  // Brenner.cxx:282-298
  ORC_REQUIRE(ij.a1->ID == C_EL);
  double s = 0;
  Atom& ai = *ij.a1;
  auto& nli = NL(ai);
  for (size_t k = 0; k < nli.size(); k++) {
    Atom& ak = (*gl_)[nli[k]];
    if (ak.ID != C_EL || &ak == ij.a2) continue;
    if (!probablyAreNeighbours(ai, ak)) continue;
    Pair ik(ai, ak, R(0, ai, ak), R(1, ai, ak));
    s += ik.f(V);
  }
  return s;
}

double Brenner::F(double x) const {  // Brenner.cxx:318-333
  if (x <= 2.0) return 1.0;
  if (x >= 3.0) return 0.0;
  return (1.0 + std::cos(M_PI * (x - 2.0))) / 2.0;
}

double Brenner::dF(double x) const {  // Brenner.cxx:335-350
  if (x <= 2.0 || x >= 3.0) return 0.0;
  return (-M_PI * std::sin(M_PI * (x - 2.0))) / 2.0;
}

double Brenner::Nconj(Pair& ij, double V) {  // Brenner.cxx:352-410
  ORC_REQUIRE(ij.a1->ID == C_EL && ij.a2->ID == C_EL);
  double N = 1.0;
  for (int side = 0; side < 2; side++) {
    Atom& c = side ? *ij.a2 : *ij.a1;
    Atom& other = side ? *ij.a1 : *ij.a2;
    auto& nlc = NL(c);
    for (size_t k = 0; k < nlc.size(); k++) {
      Atom& ak = (*gl_)[nlc[k]];
      if (&ak == &other || ak.ID != C_EL) continue;
      if (!probablyAreNeighbours(c, ak)) continue;
      Pair ki(ak, c, R(0, c, ak), R(1, c, ak));
      double f = ki.f();
      double x = Nt(ki);
      double Fx = F(x);
      if (V != 0.0) {
        ki.f(Fx * V);
        Nt(ki, f * dF(x) * V);
      }
      N += f * Fx;
    }
  }
  return N;
}

double Brenner::HN(Pair& ij, double V) {  // Brenner.cxx:412-442
  Func2O* f = nullptr;
  if (ij.a1->ID == C_EL && ij.a2->ID == C_EL) f = &hCC;
  else if (ij.a1->ID == C_EL && ij.a2->ID == H_EL) f = &hCH;
  else return 0.0;
  double nh = NH(ij), nc = NC(ij);
  double val = f->eval(nh, nc, 0);
  if (V != 0.0) {
    NH(ij, f->eval(nh, nc, 1) * V);
    NC(ij, f->eval(nh, nc, 2) * V);
  }
  return val;
}

double Brenner::FN(Pair& ij, double V) {  // Brenner.cxx:444-465
  Pair ji = ij.neg();
  if (ij.a1->ID != C_EL || ij.a2->ID != C_EL) return 0.0;
  double Ni = Nt(ij), Nj = Nt(ji), Nc = Nconj(ij);
  double val = fF.eval(Ni, Nj, Nc, 0);
  if (V != 0.0) {
    Nt(ij, fF.eval(Ni, Nj, Nc, 1) * V);
    Nt(ji, fF.eval(Ni, Nj, Nc, 2) * V);
    Nconj(ij, fF.eval(Ni, Nj, Nc, 3) * V);
  }
  return val;
}

// =================================================================== FBZL
FBZL::FBZL(double r0, double r1) : R0(r0), R1(r1) {  // FBZL.cxx:34-68
  handled = {H_EL, C_EL, Cu_EL, Au_EL, Ag_EL, Ar_EL, Xe_EL};
  nl.Rcutoff = getRcutoff();
}

bool FBZL::isHandledPair(const Atom& a, const Atom& b) const {
  bool ia = isIon(a.ID), ib = isIon(b.ID);
  if (ia && ib) return true;
  auto el = [](int id) { return id == H_EL || id == C_EL || id == Cu_EL || id == Au_EL || id == Ag_EL; };
  return (ia && el(b.ID)) || (ib && el(a.ID));
}

double FBZL::F11(Pair& ij, double V) {  // FBZL.cxx:70-99
  double R = ij.r();
  if (R > getRcutoff()) return 0.0;
  double ZA = ij.a1->Z, ZB = ij.a2->Z;
  double AS = 8.8534e-1 * AB_ / (std::pow(ZA / e_charge, 0.23) + std::pow(ZB / e_charge, 0.23));
  double Y = R / AS;
  if (V != 0.0) {
    double der = -ZA * ZB / (R * R) *
                     (0.18175 * std::exp(-3.1998 * Y) + 0.50986 * std::exp(-0.94229 * Y) +
                      0.28022 * std::exp(-0.4029 * Y) + 0.02817 * std::exp(-0.20162 * Y)) -
                 ZA * ZB / (R * AS) *
                     (0.18175 * 3.1998 * std::exp(-3.1998 * Y) + 0.50986 * 0.94229 * std::exp(-0.94229 * Y) +
                      0.28022 * 0.4029 * std::exp(-0.4029 * Y) + 0.02817 * 0.20162 * std::exp(-0.20162 * Y));
    ij.r(der * V);
  }
  return ZA * ZB / R *
         (0.18175 * std::exp(-3.1998 * Y) + 0.50986 * std::exp(-0.94229 * Y) +
          0.28022 * std::exp(-0.4029 * Y) + 0.02817 * std::exp(-0.20162 * Y));
}

double FBZL::compute(AtomsArray& atoms) {  // FBZL.cxx:101-127
  double Ei = 0;
  for (size_t i = 0; i < atoms.size(); i++) {
    Atom& a = atoms[i];
    auto& nl_ = NL(a);
    for (size_t j = 0; j < nl_.size(); j++) {
      Atom& aj = atoms[nl_[j]];
      if (a.globalIndex > aj.globalIndex) continue;
      if (!isHandledPair(a, aj)) continue;
      if (&a == &aj) continue;
      if (!(!(norm2(depos(a, aj)) > SQR(R1)))) continue;
      Pair ij(a, aj, R0, R1);
      double f = ij.f();
      double v = F11(ij, f);
      ij.f(v);
      Ei += v * f;
    }
  }
  return Ei;
}

// =================================================================== FLJ
namespace {
// ZBL universal screening (FLJ.cxx:198-224, the FBZL form): value and d/dr.
void zbl_pair(double ZA, double ZB, double r, double& v, double& d) {
  double AB_ = 0.53e-8;
  double AS = 8.8534e-1 * AB_ / (std::pow(ZA / e_charge, 0.23) + std::pow(ZB / e_charge, 0.23));
  double Y = r / AS;
  v = ZA * ZB / r *
      (0.18175 * std::exp(-3.1998 * Y) + 0.50986 * std::exp(-0.94229 * Y) + 0.28022 * std::exp(-0.4029 * Y) +
       0.02817 * std::exp(-0.20162 * Y));
  d = -ZA * ZB / (r * r) *
          (0.18175 * std::exp(-3.1998 * Y) + 0.50986 * std::exp(-0.94229 * Y) + 0.28022 * std::exp(-0.4029 * Y) +
           0.02817 * std::exp(-0.20162 * Y)) -
      ZA * ZB / (r * AS) *
          (0.18175 * 3.1998 * std::exp(-3.1998 * Y) + 0.50986 * 0.94229 * std::exp(-0.94229 * Y) +
           0.28022 * 0.4029 * std::exp(-0.4029 * Y) + 0.02817 * 0.20162 * std::exp(-0.20162 * Y));
}
void lj_pair(double sigma, double epsilon, double r, double& v, double& d) {  // FLJ.cxx:294-307
  double s_div_r = sigma / r;
  double s6 = s_div_r;
  for (int i = 2; i <= 6; i++) s6 *= s_div_r;
  double s12 = s6 * s6;
  v = 4.0 * epsilon * (s12 - s6);
  d = 4.0 * epsilon * (-12.0 * s12 / r + 6.0 * s6 / r);
}
// The cubic through (x0, v0, d0) and (x1, v1, d1) (Spline.cxx:72-98), in
// the local coordinate t = (x - x0) / (x1 - x0).
void cubic_bridge(const double x[2], const double v[2], const double d[2], double r, double& val, double& der) {
  double h = x[1] - x[0], t = (r - x[0]) / h;
  double h00 = (1 + 2 * t) * (1 - t) * (1 - t), h10 = t * (1 - t) * (1 - t);
  double h01 = t * t * (3 - 2 * t), h11 = t * t * (t - 1);
  val = h00 * v[0] + h10 * h * d[0] + h01 * v[1] + h11 * h * d[1];
  double g00 = 6 * t * (t - 1), g10 = (1 - t) * (1 - 3 * t), g01 = -6 * t * (t - 1), g11 = t * (3 * t - 2);
  der = (g00 * v[0] + g01 * v[1]) / h + g10 * d[0] + g11 * d[1];
}
}  // namespace

FLJ::FLJ(double r0, double r1) : R0(r0), R1(r1) {  // FLJ.cxx:34-156
  handled = {H_EL, C_EL, Cu_EL, Au_EL};
  nl.Rcutoff = getRcutoff();
  // Lorentz-Berthelot mixing of the like-pair parameters (FLJ.cxx:107-140)
  sigma_[Cu][Cu] = 3.05 * Ao;
  epsilon_[Cu][Cu] = 0.165565 * eV;
  sigma_[C][C] = 3.4 * Ao;
  epsilon_[C][C] = 0.002413 * eV;
  sigma_[Au][Au] = 2.569 * Ao;
  epsilon_[Au][Au] = 0.408 * eV;
  sigma_[H][H] = 2.81 * Ao;
  epsilon_[H][H] = 8.6 * 1.0 * kb;  // 8.6 K
  const int metal[2] = {Cu, Au}, hc[2] = {C, H};
  for (int m : metal)
    for (int q : hc) {
      sigma_[m][q] = sigma_[q][m] = (sigma_[m][m] + sigma_[q][q]) / 2;
      epsilon_[m][q] = epsilon_[q][m] = std::sqrt(epsilon_[m][m] * epsilon_[q][q]);
    }
  sigma_[Cu][Au] = sigma_[Au][Cu] = 0.0;  // never used: metal-metal pairs are not handled
  epsilon_[Cu][Au] = epsilon_[Au][Cu] = 0.0;
  sigma_[C][H] = sigma_[H][C] = 0.0;
  epsilon_[C][H] = epsilon_[H][C] = 0.0;
  // the bridge of every element pair (fillR_concat_, FLJ.cxx:167-253)
  const int ids[ECOUNT] = {Cu_EL, H_EL, C_EL, Au_EL};
  for (int a = 0; a < ECOUNT; a++)
    for (int b = 0; b < ECOUNT; b++) {
      Atom A1, A2;
      A1.ID = ids[a];
      A2.ID = ids[b];
      A1.setAttributesByElementID();
      A2.setAttributesByElementID();
      zbl_pair(A1.Z, A2.Z, x_[0], v_[a][b][0], d_[a][b][0]);
      lj_pair(sigma_[a][b], epsilon_[a][b], x_[1], v_[a][b][1], d_[a][b][1]);
    }
}

int FLJ::e2i(int id) const {
  switch (id) {
    case H_EL: return H;
    case C_EL: return C;
    case Cu_EL: return Cu;
    case Au_EL: return Au;
    default: throw OracleError("e2i() : unknown element");
  }
}

bool FLJ::isHandledPair(const Atom& a, const Atom& b) const {  // {H,C} x {Cu,Au}, both orders
  auto hc = [](int id) { return id == H_EL || id == C_EL; };
  auto me = [](int id) { return id == Cu_EL || id == Au_EL; };
  return (hc(a.ID) && me(b.ID)) || (me(a.ID) && hc(b.ID));
}

double FLJ::VLJ(Pair& ij) {  // FLJ.cxx:256-322
  double r = ij.r();
  int a = e2i(ij.a1->ID), b = e2i(ij.a2->ID);
  if (r < x_[0]) {
    double v, d;
    zbl_pair(ij.a1->Z, ij.a2->Z, r, v, d);
    ij.r(d);
    return v;
  }
  if (r < x_[1]) {
    double v, d;
    cubic_bridge(x_, v_[a][b], d_[a][b], r, v, d);
    ij.r(d);
    return v;
  }
  double v, d;
  lj_pair(sigma_[a][b], epsilon_[a][b], r, v, d);
  double f = ij.f();
  ij.r(d * f);
  ij.f(v);
  return f * v;
}

double FLJ::compute(AtomsArray& atoms) {  // FLJ.cxx:324-341
  double Ei = 0;
  for (size_t i = 0; i < atoms.size(); i++) {
    Atom& a = atoms[i];
    auto& nl_ = NL(a);
    for (size_t j = 0; j < nl_.size(); j++) {
      Atom& aj = atoms[nl_[j]];
      if (a.globalIndex > aj.globalIndex) continue;
      if (!isHandledPair(a, aj)) continue;
      if (&a == &aj) continue;
      if (norm2(depos(a, aj)) > SQR(R1)) continue;  // FPairwise::probablyAreNeighbours
      Pair ij(a, aj, R0, R1);
      Ei += VLJ(ij);
    }
  }
  return Ei;
}

// =================================================================== Ackland
// The published parameter tables (Ackland, Tichy, Vitek, Finnis, Phil. Mag. A
// 56 (1987) 735) as setupPotential lists them (Ackland.cxx:278-373): like
// pairs in units of the lattice constant a (eV/a^3, eV^2/a^3, knots in a),
// unlike pairs in eV/A^3 and A.
Ackland::Ackland() {
  handled = {Cu_EL, Ag_EL, Au_EL};  // Ackland.cxx:236-267
  struct Unlike {
    int a, b;
    double r[3], k[3];
  };
  const Unlike un[3] = {
      {Ag, Au, {4.40856, 3.46970, 3.00000}, {0.003566503525, 1.015075326, 0.0}},
      {Cu, Au, {4.309816, 4.047479, 3.297946}, {-0.08554551660, 0.192835880, 0.75932286}},
      {Ag, Cu, {4.15800, 3.15700, 3.00000}, {0.03533037752, 0.8518466353, 0.0}},
  };
  for (const auto& u : un)
    for (int q = 0; q < 3; q++) {
      rk_[u.a][u.b][q + 1] = rk_[u.b][u.a][q + 1] = u.r[q] * Ao;
      ak_[u.a][u.b][q + 1] = ak_[u.b][u.a][q + 1] = u.k[q] * eV / (Ao * Ao * Ao);
    }
  struct Like {
    int e;
    double lat, k[6], r[6], K[2], R[2];
  };
  const Like lk[3] = {
      {Cu, 3.615,
       {61.73525861, -108.18467800, 57.00053948, -12.88796578, 39.16381901, 0.0},
       {1.225, 1.202, 1.154, 1.050, 0.866, 0.707},
       {10.03718305, 17.06363299},
       {1.225, 0.990}},
      {Ag, 4.086,
       {20.368404, -102.36075, 94.31277, -6.220051, 31.080889, 175.56047},
       {1.2247449, 1.1547054, 1.1180065, 1.0000000, 0.8660254, 0.7071068},
       {1.458761, 42.946555},
       {1.2247449, 1.0000000}},
      {Au, 4.078,
       {29.059066, -153.14779, 148.17881, -22.20508, 72.71465, 199.26269},
       {1.2247449, 1.1547054, 1.1180065, 1.0000000, 0.8660254, 0.7071068},
       {21.930125, 284.99631},
       {1.1180065, 0.8660254}},
  };
  for (const auto& l : lk) {
    double a = l.lat * Ao;
    for (int q = 0; q < 6; q++) {
      ak_[l.e][l.e][q + 1] = l.k[q] * eV / (a * a * a);
      rk_[l.e][l.e][q + 1] = l.r[q] * a;
    }
    for (int q = 0; q < 2; q++) {
      Ak_[l.e][l.e][q + 1] = l.K[q] * (eV * eV) / (a * a * a);
      Rk_[l.e][l.e][q + 1] = l.R[q] * a;
    }
  }
  nl.Rcutoff = getRcutoff();
  // the bridge of every element pair (fillR_concat_, Ackland.cxx:377-457):
  // ZBL at 1.0 A, the knot sum at 1.4 A
  const int ids[ECOUNT] = {Cu_EL, Ag_EL, Au_EL};
  for (int a = 0; a < ECOUNT; a++)
    for (int b = 0; b < ECOUNT; b++) {
      Atom A1, A2;
      A1.ID = ids[a];
      A2.ID = ids[b];
      A1.setAttributesByElementID();
      A2.setAttributesByElementID();
      zbl_pair(A1.Z, A2.Z, x_[0], v_[a][b][0], d_[a][b][0]);
      v_[a][b][1] = knots(a, b, x_[1], d_[a][b][1]);
    }
}

int Ackland::e2i(int id) const {  // Ackland.hpp:75-84
  switch (id) {
    case Cu_EL: return Cu;
    case Ag_EL: return Ag;
    case Au_EL: return Au;
    default: throw OracleError("e2i() : unknown element");
  }
}

// sum_k ak (rk - r)^3 over the knots beyond r: 6 knots for like pairs, 3 for
// unlike (Ackland.cxx:107-128)
double Ackland::knots(int a, int b, double r, double& der) const {
  int nk = (a == b) ? 6 : 3;
  double val = 0.0;
  der = 0.0;
  for (int k = 1; k <= nk; k++) {
    double rt = rk_[a][b][k] - r;
    if (rt <= 0) continue;
    val += ak_[a][b][k] * rt * rt * rt;
    der += ak_[a][b][k] * 3.0 * rt * rt * (-1.0);
  }
  return val;
}

double Ackland::Phi(Pair& ij) {  // Ackland.cxx:59-129
  double r = ij.r();
  int a = e2i(ij.a1->ID), b = e2i(ij.a2->ID);
  if (r >= rk_[a][b][1]) return 0.0;
  double v, d;
  if (r < x_[0]) {
    zbl_pair(ij.a1->Z, ij.a2->Z, r, v, d);
  } else if (r < x_[1]) {
    cubic_bridge(x_, v_[a][b], d_[a][b], r, v, d);
  } else {
    v = knots(a, b, r, d);
  }
  ij.r(d);
  return v;
}

double Ackland::PhiCap(int a, int b, double r) const {  // Ackland.cxx:131-148
  if (r >= Rk_[a][b][1]) return 0.0;
  double val = 0.0;
  for (int k = 1; k <= 2; k++) {
    double rt = Rk_[a][b][k] - r;
    if (rt <= 0) continue;
    val += Ak_[a][b][k] * rt * rt * rt;
  }
  return val;
}

double Ackland::dPhiCap(int a, int b, double r) const {  // Ackland.cxx:150-166
  if (r >= Rk_[a][b][1]) return 0.0;
  double val = 0.0;
  for (int k = 1; k <= 2; k++) {
    double rt = Rk_[a][b][k] - r;
    if (rt <= 0) continue;
    val += Ak_[a][b][k] * 3.0 * rt * rt * (-1.0);
  }
  return val;
}

double Ackland::g(Pair& ij, double V) {  // Ackland.cxx:168-208
  double r = ij.r();
  int a = e2i(ij.a1->ID), b = e2i(ij.a2->ID);
  if (a != b) {
    double p1 = PhiCap(a, a, r), p2 = PhiCap(b, b, r);
    double s = std::sqrt(p1 * p2);
    ORC_REQUIRE(p1 * p2 >= 0);
    if (r < Rk_[a][a][1] && r < Rk_[b][b][1])
      ij.r(0.5 / s * (dPhiCap(a, a, r) * p2 + p1 * dPhiCap(b, b, r)) * V);
    return s;
  }
  if (r < Rk_[a][b][1]) ij.r(dPhiCap(a, b, r) * V);
  return PhiCap(a, b, r);
}

double Ackland::rho(Atom& ai, AtomsArray& atoms, double V) {  // Ackland.cxx:223-234
  double s = 0.0;
  for (int j : NL(ai)) {
    Pair ij(ai, atoms[j], 10.0 * Ao, 20.0 * Ao);
    s += g(ij, V);
  }
  return s;
}

double Ackland::F(Atom& ai, AtomsArray& atoms) {  // Ackland.cxx:210-221
  double r = rho(ai, atoms);
  ORC_REQUIRE(r >= 0.0);
  double Fv = -std::sqrt(r);
  if (r != 0) rho(ai, atoms, 0.5 / Fv);
  return Fv;
}

double Ackland::compute(AtomsArray& atoms) {  // Ackland.cxx:31-57
  double Ei = 0;
  for (auto& a : atoms) {
    if (!isHandled(a)) continue;
    Ei += F(a, atoms);
    for (int j : NL(a)) {
      Atom& aj = atoms[j];
      if (a.globalIndex > aj.globalIndex) continue;
      if (&a == &aj) continue;
      Pair ij(a, aj, 10.0 * Ao, 20.0 * Ao);
      Ei += Phi(ij);
    }
  }
  return Ei;
}

// =================================================================== FBM
FBM::FBM(double r0, double r1) : R0(r0), R1(r1) {  // FBM.cxx:34-53
  handled = {Cu_EL, Ar_EL};
  nl.Rcutoff = getRcutoff();
}

bool FBM::isHandledPair(const Atom& a, const Atom& b) const {  // Cu-Cu, Cu-Ar, Ar-Cu
  return (a.ID == Cu_EL && (b.ID == Cu_EL || b.ID == Ar_EL)) || (a.ID == Ar_EL && b.ID == Cu_EL);
}

void FBM::params(const Atom& a, const Atom& b, double& A, double& B) const {  // FBM.cxx:43-52
  bool ar = a.ID == Ar_EL || b.ID == Ar_EL;
  A = (ar ? 59.874 : 22.565) * 1000.0 * eV;
  B = (ar ? 72.0 : 50.88) / (10.0 * Ao);
}

double FBM::F11(Pair& ij, double V) {  // FBM.cxx:55-68
  double A, B;
  params(*ij.a1, *ij.a2, A, B);
  double R = ij.r();
  if (V != 0.0) ij.r(-A * B * std::exp(-B * R) * V);
  return A * std::exp(-B * R);
}

double FBM::compute(AtomsArray& atoms) {  // FBM.cxx:70-94
  double Ei = 0;
  for (auto& a : atoms) {
    for (int j : NL(a)) {
      Atom& aj = atoms[j];
      if (a.globalIndex > aj.globalIndex) continue;
      if (!isHandledPair(a, aj)) continue;
      if (&a == &aj) continue;
      if (norm2(depos(a, aj)) > SQR(R1)) continue;  // FPairwise::probablyAreNeighbours
      Pair ij(a, aj, R0, R1);
      double f = ij.f();
      double v = F11(ij, f);
      ij.f(v);
      Ei += v * f;
    }
  }
  return Ei;
}

// =================================================================== TightBinding
TightBinding::TightBinding() {  // TightBinding.cxx:161-189
  handled = {Cu_EL};
  alpha_ = 42.87 / (10.0 * Ao);
  beta_ = 18.00 / (10.0 * Ao);
  c_ = 12.17 * eV;
  Phi0_ = 9.892 * 1000.0 * eV;
  R_[0] = 5.0 * Ao;
  R_[1] = 5.5 * Ao;
  BM_A = 22.565 * 1000.0 * eV;
  BM_B = 50.88 / (10.0 * Ao);
  nl.Rcutoff = getRcutoff();
  // bridge ends (fillR_concat_, TightBinding.cxx:191-231): Born-Mayer at 1.0 A,
  // Phi0 exp(-alpha r) at 1.2 A
  v_[0] = BM_A * std::exp(-BM_B * x_[0]);
  d_[0] = -BM_B * BM_A * std::exp(-BM_B * x_[0]);
  v_[1] = Phi0_ * std::exp(-alpha_ * x_[1]);
  d_[1] = -alpha_ * Phi0_ * std::exp(-alpha_ * x_[1]);
}

double TightBinding::Phi(Pair& ij) {  // TightBinding.cxx:59-102
  double fvar = ij.f();
  if (fvar == 0.0) return 0.0;
  double r = ij.r();
  if (r < x_[0]) {
    ij.r(-BM_B * BM_A * std::exp(-BM_B * r));
    return BM_A * std::exp(-BM_B * r);
  }
  if (r < x_[1]) {
    double v, d;
    cubic_bridge(x_, v_, d_, r, v, d);
    ij.r(d);
    return v;
  }
  double Val = Phi0_ * std::exp(-alpha_ * r);
  ij.r(Val * (-alpha_) * fvar);
  ij.f(Val);
  return fvar * Val;
}

double TightBinding::g(Pair& ij, double V) {  // TightBinding.cxx:105-130
  double fvar = ij.f();
  if (fvar == 0.0) return 0.0;
  double r = ij.r();
  double Val = std::exp(-beta_ * r);
  ij.r(Val * (-beta_) * fvar * V);
  ij.f(Val * V);
  return fvar * Val;
}

double TightBinding::rho(Atom& ai, AtomsArray& atoms, double V) {  // TightBinding.cxx:145-159
  double s = 0.0;
  for (int j : NL(ai)) {
    Atom& aj = atoms[j];
    if (!near(ai, aj)) continue;
    Pair ij(ai, aj, R_[0], R_[1]);
    s += g(ij, V);
  }
  return s;
}

double TightBinding::F(Atom& ai, AtomsArray& atoms) {  // TightBinding.cxx:133-143
  double r = rho(ai, atoms);
  ORC_REQUIRE(r >= 0.0);
  if (r != 0) rho(ai, atoms, -c_ / (2.0 * std::sqrt(r)));
  return -c_ * std::sqrt(r);
}

double TightBinding::compute(AtomsArray& atoms) {  // TightBinding.cxx:30-56
  double Ei = 0;
  for (auto& a : atoms) {
    if (!isHandled(a)) continue;
    Ei += F(a, atoms);
    for (int j : NL(a)) {
      Atom& aj = atoms[j];
      if (a.globalIndex > aj.globalIndex) continue;
      if (&a == &aj) continue;
      if (!near(a, aj)) continue;
      Pair ij(a, aj, R_[0], R_[1]);
      Ei += Phi(ij);
    }
  }
  return Ei;
}

// =================================================================== SimLoop
void SimLoop::prepare() {  // AtomsContainer.cxx:106-123, SimLoop.cxx:236-250
  ORC_REQUIRE(!atoms.empty());
  for (size_t i = 0; i < atoms.size(); i++) {
    atoms[i].applyPBC();
    atoms[i].globalIndex = i;
    atoms[i].setAttributesByElementID();
  }
  for (auto& a : atoms) {  // fitInPBC, AtomsContainer.cxx:80-104
    if (a.PBC.x != NO_PBC_L && (a.coords.x < 0 || a.coords.x >= a.PBC.x)) throw OracleError("Atoms do not fit given PBC cell !");
    if (a.PBC.y != NO_PBC_L && (a.coords.y < 0 || a.coords.y >= a.PBC.y)) throw OracleError("Atoms do not fit given PBC cell !");
    if (a.PBC.z != NO_PBC_L && (a.coords.z < 0 || a.coords.z >= a.PBC.z)) throw OracleError("Atoms do not fit given PBC cell !");
  }
  double rc2 = fpot.getRcutoff() * 2.0;  // checkMIC, AtomsContainer.cxx:68-78
  const V3& L = arrayPBC;
  if (L.x <= rc2 || L.y <= rc2 || L.z <= rc2) throw OracleError("Rcutoff is too large for given PBC !");
  fpot.NL_init(atoms);
  fpot.NL_UpdateIfNeeded(atoms);
  prepared = true;
}

bool ThermalBath::inside(const Atom& a) const {
  switch (geom) {
    case TB_UNIVERSE:
      return true;
    case TB_BOX: {
      bool lateral = a.PBC.x != NO_PBC_L && a.PBC.y != NO_PBC_L;  // Atom::lateralPBCEnabled
      ORC_REQUIRE(lateral);
      return (a.coords.z > zMin) ||
             (lateral &&
              ((a.coords.x < 0.0 + dBoundary) || (a.coords.x > a.PBC.x - dBoundary) ||
               (a.coords.y < 0.0 + dBoundary) || (a.coords.y > a.PBC.y - dBoundary)) &&
              a.coords.z > zMinOfFreeZone);
    }
    case TB_SPHERE:
      ORC_REQUIRE(a.PBC.x == NO_PBC_L && a.PBC.y == NO_PBC_L && a.PBC.z == NO_PBC_L);
      return a.coords.z > sphereZMinOfFreeZone && norm(center - a.coords) > radius;
    default:
      return false;
  }
}

double SimLoop::bathTemperature() const {
  double e = 0.0;
  int n = 0;
  for (const auto& a : atoms) {
    if (a.fixed) continue;
    if (bathApplies(a)) {
      e += a.M * SQR(norm(a.V)) / 2.0;
      n++;
    }
  }
  if (n == 0) return 0.0;
  return e / (3.0 / 2.0 * kb * n);
}

double SimLoop::energyKin() const {  // SimLoop.cxx:500-514
  double e = 0.0;
  for (const auto& a : atoms) e += a.M * SQR(norm(a.V)) / 2.0;
  return e;
}

// One iteration of SimLoop::executeMain's loop (SimLoop.cxx:252-445), with
// the thermal bath and without file output.  A fixed dt is applied the way a
// doBeforeIteration() hook would set it.
void SimLoop::step() {
  if (!prepared) prepare();
  if (fixed_dt) dt = fixed_dt_value;
  const double Tb = bathTemperature();  // SimLoop.cxx:280, before the drift
  netForce = V3();
  double v_max = 0.0;
  for (auto& a : atoms) a.grad = V3();
  if (iteration == 0) {
    Epot = fpot.compute(atoms);
    for (auto& a : atoms) {
      if (a.fixed) continue;
      V3 force = -a.grad;
      a.an = force / a.M;
      a.an_no_tb = force / a.M;
      a.grad = V3();
    }
  }
  for (size_t j = 0; j < atoms.size(); j++) {
    Atom& a = atoms[j];
    if (a.fixed) continue;
    V3 dr = a.V * dt + a.an * dt * dt / 2.0;  // eq 1
    a.coords += dr;
    fpot.incDisplacement(j, dr);
  }
  fpot.NL_checkRequestUpdate();
  fpot.NL_UpdateIfNeeded(atoms);
  Epot = fpot.compute(atoms);
  for (auto& a : atoms) {
    V3 force = -a.grad;
    netForce += force;
    if (a.fixed) continue;
    V3 vdt2 = a.V + a.an * dt / 2.0;  // eq 2
    if (bathApplies(a)) {             // SimLoop.cxx:352-386
      double To_by_T = (std::fabs(Tb) < 1e-5) ? 0 : (bath.To / Tb);
      const double max_To_by_T = 5.0;
      if (To_by_T < -max_To_by_T) To_by_T = -max_To_by_T;
      if (To_by_T > +max_To_by_T) To_by_T = +max_To_by_T;
      V3 dforce = -a.V * a.M * bath.gamma * (1.0 - std::sqrt(To_by_T));
      V3 an_no_tb_new = force / a.M;
      V3 dv_no_tb = a.an_no_tb * dt / 2.0 + an_no_tb_new * dt / 2.0;
      a.an_no_tb = an_no_tb_new;
      V3 an_new = (force + dforce) / a.M;
      V3 dv = a.an * dt / 2.0 + an_new * dt / 2.0;
      a.an = an_new;
      double dEkin = a.M * SQR(norm(a.V + dv)) / 2.0 - a.M * SQR(norm(a.V + dv_no_tb)) / 2.0;
      energyTransferredFromBath += dEkin;
      force += dforce;
    }
    a.an = force / a.M;               // eq 3
    a.V = vdt2 + a.an * dt / 2.0;     // eq 4
    double v = norm(a.V);
    if (v > v_max && fpot.hasNeighbors(a)) v_max = v;
  }
  for (auto& a : atoms) a.applyPBC();
  simTime += dt;
  if (!fixed_dt) {  // SimLoop.cxx:409-416
    const double dt_max = 5e-16, dt_min = 1e-20;
    if (v_max != 0.0) dt = 0.05 * (1.0 * Ao) / v_max;
    else dt = dt_max;
    if (dt > dt_max) dt = dt_max;
    if (dt < dt_min) dt = dt_min;
  }
  Ekin = energyKin();
  iteration++;
}

}  // namespace orc
