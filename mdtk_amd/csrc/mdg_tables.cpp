// Host construction of the parameter and spline node tables.
//
// The numbers are the reference's parameter tables, kept identical because
// the engine must reproduce the reference's energies:
//   REBO pair/bond-order parameters   mdtk/potentials/manybody/AIREBO/REBO.cxx:856-941
//   G(cos) nodes                      AIREBO/cubici.cxx:36-95
//   P_CC, P_CH nodes                  AIREBO/cubici.cxx:112-290
//   pi_rc CC/CH/HH and T_ij nodes     AIREBO/cubici.cxx:292-536
//   AIREBO-LJ parameters              AIREBO/AIREBO_LJ.cxx:219-287
//   torsion zeta                      AIREBO/AIREBO_ETors.cxx:124-133
//   Brenner I/II parameters, H and F  Brenner/Brenner.cxx:497-611, Brenner/cubici.cxx:36-346
// The interpolants are the unique Hermite polynomials those nodes define
// (value and first derivatives, zero cross derivatives); the kernels evaluate
// them in cell-local coordinates.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mdg_device.hpp"

namespace mdg {

namespace {
// Quintic on [x1,x2] matching value, first and second derivative at both
// ends (Spline5n, mdtk/Spline5n.cxx:34-105), as local monomials in t=(x-x1)/h.
void quintic(double x1, double y1, double d1, double s1, double x2, double y2, double d2, double s2, double* c,
             double& h) {
  h = x2 - x1;
  double Y0 = y1, D0 = h * d1, S0 = h * h * s1;
  double Y1 = y2, D1 = h * d2, S1 = h * h * s2;
  c[0] = Y0;
  c[1] = D0;
  c[2] = S0 / 2.0;
  double r0 = Y1 - c[0] - c[1] - c[2];
  double r1 = D1 - c[1] - 2.0 * c[2];
  double r2 = S1 - 2.0 * c[2];
  c[3] = 10.0 * r0 - 4.0 * r1 + 0.5 * r2;
  c[4] = -15.0 * r0 + 7.0 * r1 - r2;
  c[5] = 6.0 * r0 - 3.0 * r1 + 0.5 * r2;
}

void gbuild(GSpline& g, const double* x, const double* y, const double* d, const double* s, int n) {
  g.count = n - 1;
  for (int i = 0; i < n - 1; i++) {
    g.xa[i] = x[i];
    g.xb[i] = x[i + 1];
    quintic(x[i], y[i], d[i], s[i], x[i + 1], y[i + 1], d[i + 1], s[i + 1], g.c[i], g.h[i]);
  }
}

// Node grid for one tricubic table with the reference's fill helpers:
// sym() sets F[i][j][k] and F[j][i][k]; up() copies F[i][j][kb] to all k > kb
// for (i,j) and (j,i) (cubici.cxx:292-312).
struct Grid {
  double (*t)[5][16][4];
  void sym(int i, int j, int k, double v) { t[i][j][k][0] = t[j][i][k][0] = v; }
  void symI(int i, int j, int k, double v) { t[i][j][k][1] = t[j][i][k][1] = v; }
  void symJ(int i, int j, int k, double v) { t[i][j][k][2] = t[j][i][k][2] = v; }
  void symK(int i, int j, int k, double v) { t[i][j][k][3] = t[j][i][k][3] = v; }
  void up(int ib, int jb, int kb) {
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++)
        if ((i == ib && j == jb) || (i == jb && j == ib))
          for (int k = kb + 1; k < 16; k++) t[i][j][k][0] = t[i][j][kb][0];
  }
};
// An upper bound of a tricubic Hermite interpolant with node values and
// first derivatives (cross derivatives zero) over its built cells.  Inside a
// cell the value is sum over corners of v h_a h_b h_c plus slope terms; the
// value weights are non-negative and sum to one, and each slope basis
// function is bounded by |t^3 - 2t^2 + t|, |t^3 - t^2| <= 4/27 on [0, 1]
// (the other factors lie in [0, 1]).  Outside the built cells the table is 0.
// A lower bound of a G spline over its pieces: each piece is a quintic in t
// on [0, 1], and a polynomial lies above the smallest of its Bernstein
// coefficients b_k = sum_{i<=k} C(k,i)/C(5,i) c_i.
double g_lower_bound(const GSpline& g) {
  static const double C5[6] = {1, 5, 10, 10, 5, 1};
  double lb = 1e300;
  for (int s = 0; s < g.count; s++)
    for (int k = 0; k <= 5; k++) {
      double b = 0.0, ck = 1.0;  // ck = C(k, i)
      for (int i = 0; i <= k; i++) {
        b += ck / C5[i] * g.c[s][i];
        ck = ck * (k - i) / (i + 1);
      }
      lb = std::min(lb, b);
    }
  return lb;
}

// A lower bound of a bicubic Hermite interpolant (cross derivatives zero)
// over its cells, as rc_upper_bound: the corner values' minimum less 4/27 of
// the slopes' magnitudes; beyond the table the value is 0.
double p_lower_bound(const double (*t)[5][3]) {
  double lb = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double vmin = 1e300, slope = 0.0;
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          const double* q = t[i + a][j + b];
          vmin = std::min(vmin, q[0]);
          slope += std::fabs(q[1]) + std::fabs(q[2]);
        }
      lb = std::min(lb, vmin - (4.0 / 27.0) * slope);
    }
  return lb;
}

double rc_upper_bound(const double (*t)[5][16][4], int kbuilt) {
  double ub = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < kbuilt; k++) {
        double vmax = -1e300, slope = 0.0;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
              const double* q = t[i + a][j + b][k + c];
              vmax = std::max(vmax, q[0]);
              slope += std::fabs(q[1]) + std::fabs(q[2]) + std::fabs(q[3]);
            }
        ub = std::max(ub, vmax + (4.0 / 27.0) * slope);
      }
  return ub;
}
}  // namespace

// FLJ parameters (FLJ.cxx:107-140: Lorentz-Berthelot mixing of like-pair
// values) and its ZBL-to-LJ bridge ends (fillR_concat_, FLJ.cxx:158-253).
void build_flj_table(FLJTable& t) {
  std::memset(&t, 0, sizeof(t));
  const double Ao = kAo, eV = kEV, kb = 8.617343E-5 * kEV;
  enum { Cu = 0, H = 1, C = 2, Au = 3 };
  t.sigma[Cu][Cu] = 3.05 * Ao;
  t.eps[Cu][Cu] = 0.165565 * eV;
  t.sigma[C][C] = 3.4 * Ao;
  t.eps[C][C] = 0.002413 * eV;
  t.sigma[Au][Au] = 2.569 * Ao;
  t.eps[Au][Au] = 0.408 * eV;
  t.sigma[H][H] = 2.81 * Ao;
  t.eps[H][H] = 8.6 * kb;
  for (int m : {Cu, Au})
    for (int q : {C, H}) {
      t.sigma[m][q] = t.sigma[q][m] = (t.sigma[m][m] + t.sigma[q][q]) / 2;
      t.eps[m][q] = t.eps[q][m] = std::sqrt(t.eps[m][m] * t.eps[q][q]);
    }
  t.x[0] = 1.2 * Ao;
  t.x[1] = 1.9 * Ao;
  t.R0 = 5.0 * Ao;  // common.h:49
  t.R1 = 5.5 * Ao;
  const double Z[4] = {29.0 * kECharge, 1.0 * kECharge, 6.0 * kECharge, 79.0 * kECharge};
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      // ZBL value and slope at x[0]
      double r = t.x[0];
      double AS = 8.8534e-1 * 0.53e-8 / (std::pow(Z[a] / kECharge, 0.23) + std::pow(Z[b] / kECharge, 0.23));
      double Y = r / AS;
      double e1 = std::exp(-3.1998 * Y), e2 = std::exp(-0.94229 * Y), e3 = std::exp(-0.4029 * Y),
             e4 = std::exp(-0.20162 * Y);
      double sum = 0.18175 * e1 + 0.50986 * e2 + 0.28022 * e3 + 0.02817 * e4;
      t.v[a][b][0] = Z[a] * Z[b] / r * sum;
      t.d[a][b][0] = -Z[a] * Z[b] / (r * r) * sum -
                     Z[a] * Z[b] / (r * AS) *
                         (0.18175 * 3.1998 * e1 + 0.50986 * 0.94229 * e2 + 0.28022 * 0.4029 * e3 + 0.02817 * 0.20162 * e4);
      // LJ value and slope at x[1] (no cutoff there)
      r = t.x[1];
      double s = t.sigma[a][b] / r, s6 = s;
      for (int u = 2; u <= 6; u++) s6 *= s;
      double s12 = s6 * s6;
      t.v[a][b][1] = 4.0 * t.eps[a][b] * (s12 - s6);
      t.d[a][b][1] = 4.0 * t.eps[a][b] * (-12.0 * s12 / r + 6.0 * s6 / r);
    }
}

// Ackland, Tichy, Vitek, Finnis (Phil. Mag. A 56 (1987) 735) and the unlike
// pairs of Ackland.cxx:282-299, as setupPotential sets them (Ackland.cxx:
// 278-373); the bridge ends as fillR_concat_ computes them (:377-457).
void build_eam_table(EAMTable& t) {
  std::memset(&t, 0, sizeof(t));
  const double Ao = kAo, eV = kEV;
  enum { Cu = 0, Ag = 1, Au = 2 };
  // unlike pairs: knots in A, amplitudes in eV/A^3 (the third Ag ones are 0)
  const struct { int a, b; double r[3], k[3]; } un[3] = {
      {Ag, Au, {4.40856, 3.46970, 3.00000}, {0.003566503525, 1.015075326, 0.0}},
      {Cu, Au, {4.309816, 4.047479, 3.297946}, {-0.08554551660, 0.192835880, 0.75932286}},
      {Cu, Ag, {4.15800, 3.15700, 3.00000}, {0.03533037752, 0.8518466353, 0.0}},
  };
  for (const auto& u : un)
    for (int k = 1; k <= 3; k++) {
      t.rk[u.a][u.b][k] = t.rk[u.b][u.a][k] = u.r[k - 1] * Ao;
      t.ak[u.a][u.b][k] = t.ak[u.b][u.a][k] = u.k[k - 1] * eV / (Ao * Ao * Ao);
    }
  // like pairs: lattice constant a; V knots in a, amplitudes in eV/a^3;
  // density knots in a, amplitudes in eV^2/a^3
  const struct { int e; double lat, k[6], r[6], K[2], R[2]; } lk[3] = {
      {Cu, 3.615, {61.73525861, -108.18467800, 57.00053948, -12.88796578, 39.16381901, 0.0},
       {1.225, 1.202, 1.154, 1.050, 0.866, 0.707}, {10.03718305, 17.06363299}, {1.225, 0.990}},
      {Ag, 4.086, {20.368404, -102.36075, 94.31277, -6.220051, 31.080889, 175.56047},
       {1.2247449, 1.1547054, 1.1180065, 1.0000000, 0.8660254, 0.7071068}, {1.458761, 42.946555},
       {1.2247449, 1.0000000}},
      {Au, 4.078, {29.059066, -153.14779, 148.17881, -22.20508, 72.71465, 199.26269},
       {1.2247449, 1.1547054, 1.1180065, 1.0000000, 0.8660254, 0.7071068}, {21.930125, 284.99631},
       {1.1180065, 0.8660254}},
  };
  for (const auto& l : lk) {
    const double a = l.lat * Ao, a3 = a * a * a;
    for (int k = 1; k <= 6; k++) {
      t.ak[l.e][l.e][k] = l.k[k - 1] * eV / a3;
      t.rk[l.e][l.e][k] = l.r[k - 1] * a;
    }
    for (int k = 0; k < 2; k++) {
      t.Ak[l.e][k] = l.K[k] * (eV * eV) / a3;
      t.Rk[l.e][k] = l.R[k] * a;
    }
  }
  t.x[0] = 1.0 * Ao;
  t.x[1] = 1.4 * Ao;
  const double Z[3] = {29.0 * kECharge, 47.0 * kECharge, 79.0 * kECharge};
  for (int a = 0; a < 3; a++) t.Z[a] = Z[a];
  t.rmax = 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      t.rmax = std::max(t.rmax, std::max(t.rk[a][b][1], t.Rk[a][0]));
      t.AS[a][b] = 8.8534e-1 * 0.53e-8 / (std::pow(Z[a] / kECharge, 0.23) + std::pow(Z[b] / kECharge, 0.23));
      // ZBL value and slope at x[0]
      double r = t.x[0], Y = r / t.AS[a][b];
      double e1 = std::exp(-3.1998 * Y), e2 = std::exp(-0.94229 * Y), e3 = std::exp(-0.4029 * Y),
             e4 = std::exp(-0.20162 * Y);
      double sum = 0.18175 * e1 + 0.50986 * e2 + 0.28022 * e3 + 0.02817 * e4;
      t.v[a][b][0] = Z[a] * Z[b] / r * sum;
      t.d[a][b][0] = -Z[a] * Z[b] / (r * r) * sum -
                     Z[a] * Z[b] / (r * t.AS[a][b]) *
                         (0.18175 * 3.1998 * e1 + 0.50986 * 0.94229 * e2 + 0.28022 * 0.4029 * e3 + 0.02817 * 0.20162 * e4);
      // knot sum value and slope at x[1]
      r = t.x[1];
      double val = 0, der = 0;
      for (int k = 1; k <= (a == b ? 6 : 3); k++) {
        double rt = t.rk[a][b][k] - r;
        if (rt <= 0) continue;
        val += t.ak[a][b][k] * rt * rt * rt;
        der += t.ak[a][b][k] * 3.0 * rt * rt * (-1.0);
      }
      t.v[a][b][1] = val;
      t.d[a][b][1] = der;
    }
}

// TightBinding::setupPotential (TightBinding.cxx:175-189) and the bridge
// ends of fillR_concat_ (:191-231).
void build_tb_table(TBTable& t) {
  std::memset(&t, 0, sizeof(t));
  const double Ao = kAo, eV = kEV;
  t.alpha = 42.87 / (10.0 * Ao);
  t.beta = 18.00 / (10.0 * Ao);
  t.c = 12.17 * eV;
  t.Phi0 = 9.892 * 1000.0 * eV;
  t.R0 = 5.0 * Ao;
  t.R1 = 5.5 * Ao;
  t.A = 22.565 * 1000.0 * eV;
  t.B = 50.88 / (10.0 * Ao);
  t.x[0] = 1.0 * Ao;
  t.x[1] = 1.2 * Ao;
  t.v[0] = t.A * std::exp(-t.B * t.x[0]);
  t.d[0] = -t.B * t.A * std::exp(-t.B * t.x[0]);
  t.v[1] = t.Phi0 * std::exp(-t.alpha * t.x[1]);
  t.d[1] = -t.alpha * t.Phi0 * std::exp(-t.alpha * t.x[1]);
}

// FBM::FBM (FBM.cxx:34-53) with FPairwise's default cutoff 3.1-3.5 A
// (FPairwise.hpp:34).
void build_fbm_table(FBMTable& t) {
  std::memset(&t, 0, sizeof(t));
  const double Ao = kAo, eV = kEV;
  t.A[0][0] = 22.565 * 1000.0 * eV;
  t.B[0][0] = 50.88 / (10.0 * Ao);
  t.A[0][1] = t.A[1][0] = 59.874 * 1000.0 * eV;
  t.B[0][1] = t.B[1][0] = 72.0 / (10.0 * Ao);
  t.R0 = 3.1 * Ao;
  t.R1 = 3.5 * Ao;
}

void build_tables(Tables& T, int brenner_set) {
  std::memset(&T, 0, sizeof(T));
  T.lunit = 1.0;  // the FP64 table works in the reference's cm and erg
  T.eunit = 1.0;
  const int C = 0, H = 1;
  const double Ao = kAo, eV = kEV;
  ReboParams& p = T.rebo;
  p.Q[C][C] = 0.313460 * Ao; p.Q[C][H] = p.Q[H][C] = 0.340776 * Ao; p.Q[H][H] = 0.370471 * Ao;
  p.alpha[C][C] = 4.7465391 / Ao; p.alpha[C][H] = p.alpha[H][C] = 4.1025498 / Ao; p.alpha[H][H] = 3.5362986 / Ao;
  p.A[C][C] = 10953.544 * eV; p.A[C][H] = p.A[H][C] = 149.94099 * eV; p.A[H][H] = 32.817356 * eV;
  p.B[0][C][C] = 12388.792 * eV; p.B[0][C][H] = p.B[0][H][C] = 32.355187 * eV; p.B[0][H][H] = 29.632593 * eV;
  p.B[1][C][C] = 17.567065 * eV;
  p.B[2][C][C] = 30.714932 * eV;
  p.beta[0][C][C] = 4.7204523 / Ao; p.beta[0][C][H] = p.beta[0][H][C] = 1.4344581 / Ao;
  p.beta[0][H][H] = 1.7158922 / Ao;
  p.beta[1][C][C] = 1.4332132 / Ao;
  p.beta[2][C][C] = 1.3826913 / Ao;
  p.rho[C][C] = 0.0; p.rho[C][H] = p.rho[H][C] = 1.09 * Ao; p.rho[H][H] = 0.7415887 * Ao;
  p.R[C][C][0] = 1.7 * Ao; p.R[H][H][0] = 1.1 * Ao; p.R[C][H][0] = p.R[H][C][0] = 1.3 * Ao;
  p.R[C][C][1] = 2.0 * Ao; p.R[H][H][1] = 1.7 * Ao; p.R[C][H][1] = p.R[H][C][1] = 1.8 * Ao;
  p.Rp[C][C][0] = 1.7 * Ao; p.Rp[H][H][0] = 1.1 * Ao; p.Rp[C][H][0] = p.Rp[H][C][0] = 1.3 * Ao;
  p.Rp[C][C][1] = 2.0 * Ao; p.Rp[H][H][1] = 1.7 * Ao; p.Rp[C][H][1] = p.Rp[H][C][1] = 1.6 * Ao;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) p.R2sq[a][b] = p.R[a][b][1] * p.R[a][b][1];
  p.Ntot[0] = 2;
  p.Ntot[1] = 3;
  p.expscale = 4.0 / Ao;  // REBO.cxx:221
  {
    const double x1[5] = {-1.0, -2.0 / 3.0, -1.0 / 2.0, -1.0 / 3.0, 1.0};
    const double y1[5] = {-0.010000, 0.028207, 0.052804, 0.097321, 1.000000};
    const double d1[5] = {0.104000, 0.131443, 0.170000, 0.400000, 2.834570};
    const double s1[5] = {0.000000, 0.140229, 0.370000, 1.980000, 10.264700};
    gbuild(p.gC1, x1, y1, d1, s1, 5);
    const double y2[5] = {-0.010000, 0.028207, 0.052804, 0.097321, 8.000000};
    const double d2[5] = {0.104000, 0.131443, 0.170000, 0.400000, 20.243600};
    const double s2[5] = {0.000000, 0.140229, 0.370000, 1.980000, 43.9336000};
    gbuild(p.gC2, x1, y2, d2, s2, 5);
    const double xh[4] = {-1.0, -5.0 / 6.0, -1.0 / 2.0, 1.0};
    const double yh[4] = {11.235700, 12.595300, 16.811100, 19.991800};
    const double dh[4] = {0.000000, 13.854300, 8.641230, 0.333013};
    const double sh[4] = {115.115000, 32.361800, -25.061700, -0.474189};
    gbuild(p.gH, xh, yh, dh, sh, 4);
  }
  // P_CC (zero derivatives) and P_CH
  {
    auto& q = T.P[0];
    q[2][0][0] = -0.000500; q[3][0][0] = 0.016125; q[1][1][0] = -0.010960;
    q[2][1][0] = 0.006326; q[0][2][0] = -0.027603; q[1][2][0] = 0.003180;
    auto& w = T.P[1];
    w[1][0][0] = 0.209337; w[2][0][0] = -0.064450; w[3][0][0] = -0.303928; w[0][1][0] = 0.010000;
    w[1][1][0] = -0.125123; w[2][1][0] = -0.298905; w[0][2][0] = -0.122042; w[1][2][0] = -0.300529;
    w[0][3][0] = -0.307585;
  }
  {
    Grid g{T.T3[0]};  // pi_rc CC
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
  {
    Grid g{T.T3[1]};  // pi_rc CH
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
  {
    Grid g{T.T3[2]};  // pi_rc HH
    g.sym(1, 1, 1, 0.124916);
  }
  {
    Grid g{T.T3[3]};  // T_ij
    g.sym(2, 2, 1, -0.035140);
    g.sym(2, 2, 2, -0.004048); g.up(2, 2, 2);
  }
  // pi_rc upper bounds in (i, j) orientation; H-C evaluates F_CH with the
  // arguments swapped, which the bound over the whole table covers.  The
  // margin covers the FP32 form's rounding of the evaluation.
  {
    const double margin = 1e-5;
    T.rcmax[C][C] = rc_upper_bound(T.T3[0], 15) + margin;
    T.rcmax[C][H] = T.rcmax[H][C] = rc_upper_bound(T.T3[1], 15) + margin;
    T.rcmax[H][H] = rc_upper_bound(T.T3[2], 15) + margin;
    // lower bounds of G and P for the per-atom b* bound; the margins cover
    // the FP32 evaluation (|G| <= 20, |P| < 1)
    T.glb[0] = std::min(g_lower_bound(T.rebo.gC1), g_lower_bound(T.rebo.gC2)) - 1e-4;
    T.glb[1] = g_lower_bound(T.rebo.gH) - 1e-3;
    T.plb[C] = p_lower_bound(T.P[0]) - 1e-4;
    T.plb[H] = p_lower_bound(T.P[1]) - 1e-4;
  }
  // AIREBO-LJ
  LJParams& l = T.lj;
  l.sigma[C][C] = 3.40 * Ao;
  l.sigma[H][H] = 2.65 * Ao;
  l.sigma[C][H] = l.sigma[H][C] = 0.5 * (l.sigma[C][C] + l.sigma[H][H]);
  l.zeta[C][C] = 0.00284 * eV;
  l.zeta[H][H] = 0.00150 * eV;
  l.zeta[C][H] = l.zeta[H][C] = std::sqrt(l.zeta[C][C] * l.zeta[H][H]);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      l.RLJ[a][b][0] = l.sigma[a][b];
      l.RLJ[a][b][1] = l.sigma[a][b] * std::pow(2.0, 1.0 / 6.0);
    }
  l.R[C][C][0] = 5.0 * Ao; l.R[H][H][0] = 4.0 * Ao; l.R[C][H][0] = l.R[H][C][0] = 4.5 * Ao;
  l.R[C][C][1] = l.R[C][C][0] + 0.5 * Ao;
  l.R[H][H][1] = l.R[H][H][0] + 0.5 * Ao;
  l.R[C][H][1] = l.R[H][C][1] = l.R[C][H][0] + 0.5 * Ao;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) l.R2sq[a][b] = l.R[a][b][1] * l.R[a][b][1];
  l.b[C][C][0] = 0.77; l.b[H][H][0] = 0.32; l.b[C][H][0] = l.b[H][C][0] = 0.75;
  l.b[C][C][1] = 0.81; l.b[H][H][1] = 0.42; l.b[C][H][1] = l.b[H][C][1] = 0.90;
  T.zetaT[C][C] = 0.3079 * eV;
  T.zetaT[H][H] = 0.1250 * eV;
  T.zetaT[C][H] = T.zetaT[H][C] = 0.1787 * eV;

  // Brenner
  BrennerParams& b = T.bren;
  if (brenner_set == 0) {
    b.Re[C][C] = 1.315 * Ao; b.De[C][C] = 6.325 * eV; b.beta[C][C] = 1.5 / Ao; b.S[C][C] = 1.29;
    b.delta[C][C] = 0.80469; b.delta[H][H] = 0.80469;
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++)
        for (int z = 0; z < 2; z++) b.alpha[x][y][z] = 3.0 / Ao;
    b.ao = 0.011304; b.co = 19.0; b.do_ = 2.5; b.G_HH = 4.0;
    b.S[C][H] = 1.7386;
  } else {
    b.Re[C][C] = 1.39 * Ao; b.De[C][C] = 6.0 * eV; b.beta[C][C] = 2.1 / Ao; b.S[C][C] = 1.22;
    b.delta[C][C] = 0.5; b.delta[H][H] = 0.5;
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++)
        for (int z = 0; z < 2; z++) b.alpha[x][y][z] = 4.0 / Ao;
    b.ao = 0.00020813; b.co = 330.0; b.do_ = 3.5; b.G_HH = 12.33;
    b.S[C][H] = 1.69077;
  }
  b.alpha[C][H][C] = 0.0; b.alpha[C][C][H] = 0.0; b.alpha[C][C][C] = 0.0;
  b.Re[H][H] = 0.74144 * Ao; b.Re[C][H] = b.Re[H][C] = 1.1199 * Ao;
  b.De[H][H] = 4.7509 * eV; b.De[C][H] = b.De[H][C] = 3.6422 * eV;
  b.beta[H][H] = 1.9436 / Ao; b.beta[C][H] = b.beta[H][C] = 1.9583 / Ao;
  b.S[H][H] = 2.3432; b.S[H][C] = b.S[C][H];
  b.R[C][C][0] = 1.7 * Ao; b.R[H][H][0] = 1.1 * Ao; b.R[C][H][0] = b.R[H][C][0] = 1.3 * Ao;
  b.R[C][C][1] = 2.0 * Ao; b.R[H][H][1] = 1.7 * Ao; b.R[C][H][1] = b.R[H][C][1] = 1.8 * Ao;
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++) b.R2sq[x][y] = b.R[x][y][1] * b.R[x][y][1];
  {
    auto& hc = T.BH[0];
    auto& hh = T.BH[1];
    if (brenner_set == 0) {
      hc[1][1][0] = -0.0175; hc[2][0][0] = -0.0070; hc[3][0][0] = 0.0119; hc[1][2][0] = 0.0115; hc[2][1][0] = 0.0118;
      hh[1][0][0] = -0.0760; hh[2][0][0] = -0.2163; hh[3][0][0] = -0.3375; hh[0][1][0] = -0.1792;
      hh[0][2][0] = -0.2407; hh[1][1][0] = -0.2477; hh[2][1][0] = -0.3320; hh[0][3][0] = -0.3323;
      hh[1][2][0] = -0.3321;
      hh[2][0][1] = -0.13075; hh[1][1][1] = -0.0764; hh[0][2][2] = -0.07655; hh[1][1][2] = -0.12805;
    } else {
      hc[1][1][0] = -0.0226; hc[2][0][0] = -0.0061; hc[3][0][0] = 0.0173; hc[1][2][0] = 0.0149; hc[2][1][0] = 0.0160;
      hh[1][0][0] = -0.0984; hh[2][0][0] = -0.2878; hh[3][0][0] = -0.4507; hh[0][1][0] = -0.2479;
      hh[0][2][0] = -0.3221; hh[1][1][0] = -0.3344; hh[2][1][0] = -0.4438; hh[0][3][0] = -0.4460;
      hh[1][2][0] = -0.4449;
      hh[2][0][1] = -0.17615; hh[1][1][1] = -0.09795; hh[0][2][2] = -0.09905; hh[1][1][2] = -0.17325;
    }
    auto& F = T.BF;
    auto v = [&](int i, int j, int k, double x) { F[i][j][k][0] = x; };
    auto di = [&](int i, int j, int k, double x) { F[i][j][k][1] = x; };
    auto dj = [&](int i, int j, int k, double x) { F[i][j][k][2] = x; };
    if (brenner_set == 0) {
      v(2, 3, 1, -0.0465); v(3, 2, 1, -0.0465); v(2, 3, 2, -0.0465); v(3, 2, 2, -0.0465);
      v(1, 2, 2, -0.0355); v(2, 1, 2, -0.0355); v(1, 2, 1, 0.0126); v(2, 1, 1, 0.0126);
      v(1, 3, 1, -0.1130); v(3, 1, 1, -0.1130); v(1, 3, 2, -0.1130); v(3, 1, 2, -0.1130);
      v(0, 3, 1, -0.1220); v(3, 0, 1, -0.1220); v(0, 3, 2, -0.1220); v(3, 0, 2, -0.1220);
      v(0, 2, 2, -0.0445); v(2, 0, 2, -0.0445); v(0, 2, 1, 0.0320); v(2, 0, 1, 0.0320);
      v(0, 1, 1, 0.1100); v(1, 0, 1, 0.1100);
      v(1, 1, 2, 0.0074); v(1, 1, 1, 0.1511); v(2, 2, 1, 0.0750);
      di(2, 0, 1, -0.1160); di(2, 1, 1, -0.13205); di(2, 0, 2, -0.0610); di(1, 2, 2, 0.02225);
      di(1, 3, 2, 0.03775); di(2, 3, 2, 0.0565); di(2, 3, 1, 0.0565); di(2, 1, 2, -0.0602);
      dj(0, 2, 1, -0.1160); dj(1, 2, 1, -0.13205); dj(0, 2, 2, -0.0610); dj(2, 1, 2, 0.02225);
      dj(3, 1, 2, 0.03775); dj(3, 2, 2, 0.0565); dj(3, 2, 1, 0.0565); dj(1, 2, 2, -0.0602);
    } else {
      v(2, 3, 1, -0.0363); v(3, 2, 1, -0.0363); v(2, 3, 2, -0.0363); v(3, 2, 2, -0.0363);
      v(1, 2, 2, -0.0243); v(2, 1, 2, -0.0243); v(1, 2, 1, 0.0120); v(2, 1, 1, 0.0120);
      v(1, 3, 1, -0.0903); v(3, 1, 1, -0.0903); v(1, 3, 2, -0.0903); v(3, 1, 2, -0.0903);
      v(0, 3, 1, -0.0904); v(3, 0, 1, -0.0904); v(0, 3, 2, -0.0904); v(3, 0, 2, -0.0904);
      v(0, 2, 2, -0.0269); v(2, 0, 2, -0.0269); v(0, 2, 1, 0.0427); v(2, 0, 1, 0.0427);
      v(0, 1, 1, 0.0996); v(1, 0, 1, 0.0996);
      v(1, 1, 2, 0.0108); v(1, 1, 1, 0.1264); v(2, 2, 1, 0.0605);
      di(2, 0, 1, -0.0950); di(2, 1, 1, -0.10835); di(2, 0, 2, -0.0452); di(1, 2, 2, 0.01345);
      di(1, 3, 2, 0.02705); di(2, 3, 2, 0.04515); di(2, 3, 1, 0.04515); di(2, 1, 2, -0.05055);
      dj(0, 2, 1, -0.0950); dj(1, 2, 1, -0.10835); dj(0, 2, 2, -0.0452); dj(2, 1, 2, 0.01345);
      dj(3, 1, 2, 0.02705); dj(3, 2, 2, 0.04515); dj(3, 2, 1, 0.04515); dj(1, 2, 2, -0.05055);
    }
    for (int i = 0; i < 5; i++)  // F[i][j][k>=3] = F[i][j][2]
      for (int j = 0; j < 5; j++)
        for (int k = 3; k < 6; k++) F[i][j][k][0] = F[i][j][2][0];
  }
}

// The FP32 table of the mixed-precision form: lengths in Angstrom, energies in
// eV, every entry rounded once from the FP64 CGS table.
void convert_tables(const Tables& d, TablesT<float>& f) {
  std::memset(&f, 0, sizeof(f));
  const double L = 1.0 / kAo;  // cm -> Angstrom
  const double E = 1.0 / kEV;  // erg -> eV
  f.lunit = kAo;
  f.eunit = kEV;
  auto cv = [](float* out, const double* in, int n, double s) {
    for (int q = 0; q < n; q++) out[q] = float(in[q] * s);
  };
  const ReboParams& a = d.rebo;
  ReboParamsT<float>& b = f.rebo;
  cv(&b.Q[0][0], &a.Q[0][0], 4, L);
  cv(&b.alpha[0][0], &a.alpha[0][0], 4, 1.0 / L);
  cv(&b.A[0][0], &a.A[0][0], 4, E);
  cv(&b.B[0][0][0], &a.B[0][0][0], 12, E);
  cv(&b.beta[0][0][0], &a.beta[0][0][0], 12, 1.0 / L);
  cv(&b.rho[0][0], &a.rho[0][0], 4, L);
  cv(&b.R[0][0][0], &a.R[0][0][0], 8, L);
  cv(&b.Rp[0][0][0], &a.Rp[0][0][0], 8, L);
  cv(&b.R2sq[0][0], &a.R2sq[0][0], 4, L * L);
  cv(b.Ntot, a.Ntot, 2, 1.0);
  b.expscale = float(a.expscale / L);
  const GSpline* gs[3] = {&a.gC1, &a.gC2, &a.gH};
  GSplineT<float>* gd[3] = {&b.gC1, &b.gC2, &b.gH};
  for (int g = 0; g < 3; g++) {
    gd[g]->count = gs[g]->count;
    cv(gd[g]->xa, gs[g]->xa, 4, 1.0);
    cv(gd[g]->xb, gs[g]->xb, 4, 1.0);
    cv(gd[g]->h, gs[g]->h, 4, 1.0);
    cv(&gd[g]->c[0][0], &gs[g]->c[0][0], 24, 1.0);
  }
  cv(&f.lj.sigma[0][0], &d.lj.sigma[0][0], 4, L);
  cv(&f.lj.zeta[0][0], &d.lj.zeta[0][0], 4, E);
  cv(&f.lj.RLJ[0][0][0], &d.lj.RLJ[0][0][0], 8, L);
  cv(&f.lj.R[0][0][0], &d.lj.R[0][0][0], 8, L);
  cv(&f.lj.R2sq[0][0], &d.lj.R2sq[0][0], 4, L * L);
  cv(&f.lj.b[0][0][0], &d.lj.b[0][0][0], 8, 1.0);
  cv(&f.zetaT[0][0], &d.zetaT[0][0], 4, E);
  cv(&f.P[0][0][0][0], &d.P[0][0][0][0], 2 * 5 * 5 * 3, 1.0);
  cv(&f.T3[0][0][0][0][0], &d.T3[0][0][0][0][0], 4 * 5 * 5 * 16 * 4, 1.0);
  std::memcpy(f.rcmax, d.rcmax, sizeof(f.rcmax));
  std::memcpy(f.glb, d.glb, sizeof(f.glb));
  std::memcpy(f.plb, d.plb, sizeof(f.plb));
}

}  // namespace mdg
