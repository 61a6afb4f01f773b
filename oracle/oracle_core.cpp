// Oracle: CPU restatement of the MDTK force-and-integration path.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).  Built with
// -ffp-contract=off so that a*b+c rounds twice, as in the reference build.
#include "oracle_core.hpp"

#include <algorithm>
#include <cstring>

namespace orc {

// ------------------------------------------------------------------ atoms
void Atom::setAttributesByElementID() {  // mdtk/Atom.hpp:98-113
  switch (ID) {
    case H_EL: Z = 1.0 * e_charge; M = 1.0 * amu; break;
    case C_EL: Z = 6.0 * e_charge; M = 12.0 * amu; break;
    case Cu_EL: Z = 29.0 * e_charge; M = 63.546 * amu; break;
    case Ag_EL: Z = 47.0 * e_charge; M = 107.868 * amu; break;
    case Au_EL: Z = 79.0 * e_charge; M = 196.967 * amu; break;
    case Ar_EL: Z = 18.0 * e_charge; M = 39.948 * amu; break;
    case Xe_EL: Z = 54.0 * e_charge; M = 131.293 * amu; break;
    default: throw OracleError("unknown element");
  }
}

static void wrap_axis(double& c, double L, int& cnt) {
  if (L != NO_PBC_L) {
    while (c < 0) { c += L; --cnt; }
    while (c >= L) { c -= L; ++cnt; }
  } else {
    ORC_REQUIRE(cnt == 0);
  }
}

void Atom::applyPBC() {  // mdtk/Atom.cxx:93-125
  wrap_axis(coords.x, PBC.x, pbc_count[0]);
  wrap_axis(coords.y, PBC.y, pbc_count[1]);
  wrap_axis(coords.z, PBC.z, pbc_count[2]);
}

// Minimum-image displacement a1-a2, mdtk/Atom.hpp:115-140: unwrap an atom
// whose axis is periodic against one whose axis is not, then a single +-L
// shift when both share the same period and |r| > L/2.
static inline double depos_axis(double r, double L1, double L2, int c1, int c2) {
  if (L1 == NO_PBC_L && L2 != NO_PBC_L) r -= L2 * c2;
  if (L2 == NO_PBC_L && L1 != NO_PBC_L) r += L1 * c1;
  if (L1 == L2) {
    if (L1 != NO_PBC_L && std::fabs(r) > L1 * 0.5) r += (r > 0) ? (-L1) : (L1);
  }
  return r;
}

V3 depos(const Atom& a1, const Atom& a2) {
  V3 r = a1.coords - a2.coords;
  r.x = depos_axis(r.x, a1.PBC.x, a2.PBC.x, a1.pbc_count[0], a2.pbc_count[0]);
  r.y = depos_axis(r.y, a1.PBC.y, a2.PBC.y, a1.pbc_count[1], a2.pbc_count[1]);
  r.z = depos_axis(r.z, a1.PBC.z, a2.PBC.z, a1.pbc_count[2], a2.pbc_count[2]);
  return r;
}

// ------------------------------------------------------------------ pairs
Pair::Pair(Atom& i, Atom& j, bool init) : a1(&i), a2(&j) {  // AtomsPair.hpp:117-131
  if (init) {
    rv = depos(i, j);
    r_ = norm(rv);
    r2_ = r_ * r_;
    u = rv / r_;
    f_ = 1.0;
    df_ = 0.0;
  }
}

Pair::Pair(Atom& i, Atom& j, double R1, double R2, double alter)  // AtomsPair.hpp:132-156
    : a1(&i), a2(&j) {
  rv = depos(i, j);
  r_ = norm(rv);
  r2_ = r_ * r_;
  u = rv / r_;
  ralt_ = alter;
  alt = (alter != 0.0);
  if (r_ < R1) {
    f_ = 1.0;
  } else if (r_ < R2) {
    f_ = (1.0 + std::cos(M_PI * (r_ - R1) / (R2 - R1))) / 2.0;
    df_ = (-(M_PI / (R2 - R1)) * std::sin(M_PI * (r_ - R1) / (R2 - R1))) / 2.0;
  }
}

Pair Pair::neg() const {  // AtomsPair.hpp:103-116
  Pair p(*a2, *a1, false);
  p.rv = -rv;
  p.r_ = r_;
  p.r2_ = r2_;
  p.u = -u;
  p.f_ = f_;
  p.df_ = df_;
  p.ralt_ = ralt_;
  p.alt = alt;
  return p;
}

// The reference's r_alter derivative: the unit vector of the *actual*
// separation, even though the value is the constant r_alter
// (AtomsPair.hpp:53-63).
bool g_alter_exact = false;
bool g_vir_on = false;
double g_vir[6] = {0, 0, 0, 0, 0, 0};

V3 Pair::dr_alter(const Atom& d) const {
  if (!alt) return dr(d);
  if (g_alter_exact) return V3();  // test switch: exact derivative of the constant
  V3 t = (rv / norm(rv)) * ralt_;
  double m = norm(t);
  if (a1 == &d) return t / m;
  if (a2 == &d) return -(t / m);
  return V3();
}

double Pair::r_alter(double V) {  // AtomsPair.hpp:44-52
  if (!alt) return r(V);
  if (V != 0.0) {
    a1->grad += dr_alter(*a1) * V;
    a2->grad += dr_alter(*a2) * V;
    if (g_vir_on) vir_push(rv, dr_alter(*a1) * V);
  }
  return ralt_;
}

// cos(theta_jik) for ij=(i,j), ik=(i,k), FGeneral.hpp:138-185.  Adjoint in
// vector form: d cos/d a = b/(|a||b|) - cos a/|a|^2 with a=x_i-x_j, b=x_i-x_k.
double CosTheta(Pair& ij, Pair& ik, double V) {
  double e1 = dot(ij.rv, ik.rv);
  double e2 = ij.r_ * ik.r_;
  double c = e1 / e2;
  ORC_REQUIRE(c >= -1.0 - 1e-7 && c <= 1.0 + 1e-7);
  double cc = c;
  if (cc < -1.0) cc = -1.0;
  if (cc > 1.0) cc = 1.0;
  if (V != 0.0) {
    V3 ga = ik.rv / e2 - ij.rv * (c / (ij.r_ * ij.r_));
    V3 gb = ij.rv / e2 - ik.rv * (c / (ik.r_ * ik.r_));
    ij.a1->grad += (ga + gb) * V;
    ij.a2->grad += (-ga) * V;
    ik.a2->grad += (-gb) * V;
    if (g_vir_on) {
      vir_push(ij.rv, ga * V);
      vir_push(ik.rv, gb * V);
    }
  }
  return cc;
}

double SinTheta(Pair& ij, Pair& ik) {  // FGeneral.hpp:187-199
  double s = norm(cross(ij.rv, ik.rv)) / (ij.r_ * ik.r_);
  ORC_REQUIRE(s >= -1.0 - 1e-7 && s <= 1.0 + 1e-7);
  if (s < -1.0) s = -1.0;
  if (s > 1.0) s = 1.0;
  return s;
}

// Cosine of the angle between planes (j,i,k) and (i,j,l); FGeneral.hpp:201-305.
// P = (-a) x b, Q = a x c with a = x_i-x_j, b = x_i-x_k, c = x_j-x_l.
// Adjoint: gP = Q/(|P||Q|) - cos P/|P|^2, gQ likewise;
// d/da = gP x b + c x gQ, d/db = a x gP, d/dc = gQ x a.
double CosDihedral(Pair& ij, Pair& ik, Pair& jl, double V) {
  const V3& a = ij.rv;
  const V3& b = ik.rv;
  const V3& c = jl.rv;
  V3 P = cross(-a, b);
  V3 Q = cross(a, c);
  double e1 = dot(P, Q);
  double p2 = norm2(P), q2 = norm2(Q);
  double e2 = std::sqrt(p2) * std::sqrt(q2);
  double cs = (e2 == 0.0) ? 0.0 : e1 / e2;
  double cc = cs;
  if (cc < -1.0) cc = -1.0;
  if (cc > 1.0) cc = 1.0;
  if (V != 0.0 && e2 != 0.0) {
    V3 gP = Q / e2 - P * (cs / p2);
    V3 gQ = P / e2 - Q * (cs / q2);
    V3 da = cross(gP, b) + cross(c, gQ);
    V3 db = cross(a, gP);
    V3 dc = cross(gQ, a);
    ij.a1->grad += (da + db) * V;
    ij.a2->grad += (dc - da) * V;
    ik.a2->grad += (-db) * V;
    jl.a2->grad += (-dc) * V;
    if (g_vir_on) {
      vir_push(a, da * V);
      vir_push(b, db * V);
      vir_push(c, dc * V);
    }
  }
  return cc;
}

// ------------------------------------------------------------------ NL
bool NeighbourList::movedTooMuch() const {  // NeighbourList.cxx:108-133
  ORC_REQUIRE(Rcutoff > 0.0);
  double d1 = 0.0, d2 = 0.0;
  for (const V3& d : displacements) {
    double m = norm(d);
    if (m >= d1) {
      d2 = d1;
      d1 = m;
    } else if (m >= d2) {
      d2 = m;
    }
  }
  return (d1 + d2 > 0.44 * Rcutoff);
}

// NeighbourList.cxx:31-106: all-pairs pass; pair (i,j) enters list p iff
// both atoms are handled by p and d^2 < ((1+0.44) Rc_p)^2.  Lists come out in
// ascending index order.
void nl_update(AtomsArray& atoms, std::vector<Potential*>& pots) {
  if (pots.empty()) return;
  size_t n = atoms.size();
  for (auto* p : pots) {
    p->nl.nl.assign(n, {});
    p->nl.displacements.assign(n, V3());
  }
  double rmax2 = 0;
  for (auto* p : pots) rmax2 = std::max(rmax2, SQR((1.0 + 0.44) * p->nl.Rcutoff));
  ORC_REQUIRE(rmax2 > 0.0);
  for (size_t i = 0; i < n; i++) {
    for (size_t j = i + 1; j < n; j++) {
      double d2 = norm2(depos(atoms[i], atoms[j]));
      if (d2 > rmax2) continue;
      for (auto* p : pots) {
        if (!p->isHandled(atoms[i]) || !p->isHandled(atoms[j])) continue;
        if (d2 < SQR((1.0 + 0.44) * p->nl.Rcutoff)) {
          p->nl.nl[i].push_back((int)j);
          p->nl.nl[j].push_back((int)i);
        }
      }
    }
  }
}

void Proxy::NL_UpdateIfNeeded(AtomsArray& atoms) {  // FProxy.hpp:70-83
  std::vector<Potential*> upd;
  for (auto* p : pots)
    if (p->nl.updateRequested) {
      upd.push_back(p);
      p->nl.updateRequested = false;
    }
  nl_update(atoms, upd);
}

double Proxy::compute(AtomsArray& atoms) {  // FProxy.cxx:45-61
  for (auto& a : atoms) a.grad = V3();
  double E = 0;
  lastE.assign(pots.size(), 0.0);
  for (size_t i = 0; i < pots.size(); i++) {
    double e = pots[i]->compute(atoms);
    lastE[i] = e;
    E += e;
  }
  return E;
}

// ------------------------------------------------------------------ splines
// Dense LU with partial pivoting (row of largest |a| in the column), then
// forward/back substitution.  Used to solve each cell's interpolation system.
static void lu_solve(std::vector<double>& A, std::vector<double>& b, int n) {
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A[i * n + k]) > best) {
        best = std::fabs(A[i * n + k]);
        p = i;
      }
    if (best == 0.0) throw OracleError("singular spline system");
    if (p != k) {
      for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; i++) {
      double l = A[i * n + k] / A[k * n + k];
      A[i * n + k] = l;
      for (int j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) b[i] -= A[i * n + j] * b[j];
  for (int i = n - 1; i >= 0; i--) {
    for (int j = i + 1; j < n; j++) b[i] -= A[i * n + j] * b[j];
    b[i] /= A[i * n + i];
  }
}

static inline double ipow(double a, int n) {  // mdtk/tools.hpp:107-135
  double v = 1.0;
  for (int i = 0; i < n; i++) v *= a;
  return v;
}

// Monomial enumeration order of SplineMultiD's constructor
// (mdtk/SplineAux.hpp:178-206): first coordinate's power runs fastest.
void LocalPoly::setup(int d, int o) {
  dim = d;
  order = o;
  int nt = 1;
  for (int i = 0; i < d; i++) nt *= (o + 1);
  pw.assign(nt * d, 0);
  for (int t = 0; t < nt; t++) {
    int r = t;
    for (int i = 0; i < d; i++) {
      pw[t * d + i] = r % (o + 1);
      r /= (o + 1);
    }
  }
  c.assign(nt, 0.0);
}

// Term-by-term sum in term order (SplineAux.hpp:80-94,230-241), local coords.
double LocalPoly::eval(const double* x, int dx) const {
  if (!built) return 0.0;
  int nt = (int)c.size();
  double val = 0;
  for (int t = 0; t < nt; t++) {
    double mult = 1.0;
    double v;
    if (dx >= 0) {
      mult = pw[t * dim + dx];
      if (mult == 0.0 || c[t] == 0.0) continue;
    } else if (c[t] == 0.0) {
      continue;
    }
    v = mult * c[t];
    for (int i = 0; i < dim; i++) {
      int p = pw[t * dim + i] - (i == dx ? 1 : 0);
      v *= ipow(x[i] - x0[i], p);
    }
    val += v;
  }
  return val;
}

// Row of the design matrix for a (possibly differentiated) monomial basis at
// local point xl: d^{ord} / dx_{ds}... evaluated at xl (woC, SplineAux.hpp:96-110).
static void design_row(const LocalPoly& p, const double* xl, const int* dord, double* row) {
  int nt = (int)p.c.size();
  for (int t = 0; t < nt; t++) {
    double v = 1.0;
    for (int i = 0; i < p.dim; i++) {
      int pw = p.pw[t * p.dim + i];
      for (int k = 0; k < dord[i]; k++) {
        v *= pw;
        pw -= 1;
      }
      if (v == 0.0) break;
      v *= ipow(xl[i], pw);
    }
    row[t] = v;
  }
}

void Spline5nO::init(double x1, double y1, double dy1, double d2y1, double x2, double y2, double dy2,
                     double d2y2) {
  p.setup(1, 5);
  p.x0[0] = x1;
  xa = x1;
  xb = x2;
  const double xs[2] = {0.0, x2 - x1};
  const double vs[3][2] = {{y1, y2}, {dy1, dy2}, {d2y1, d2y2}};
  std::vector<double> A(36), b(6);
  int e = 0;
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 2; k++) {
      int dord[1] = {d};
      design_row(p, &xs[k], dord, &A[e * 6]);
      b[e] = vs[d][k];
      e++;
    }
  lu_solve(A, b, 6);
  p.c = b;
  p.built = true;
}

void Spline2DO::init(double x0, double y0, const double z[2][2], const double dzdx[2][2],
                     const double dzdy[2][2]) {
  p.setup(2, 3);
  p.x0[0] = x0;
  p.x0[1] = y0;
  std::vector<double> A(256), b(16);
  int e = 0;
  const int ords[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
  for (int q = 0; q < 4; q++)
    for (int i1 = 0; i1 < 2; i1++)
      for (int i2 = 0; i2 < 2; i2++) {
        double xl[2] = {(double)i1, (double)i2};
        design_row(p, xl, ords[q], &A[e * 16]);
        b[e] = q == 0 ? z[i1][i2] : q == 1 ? dzdx[i1][i2] : q == 2 ? dzdy[i1][i2] : 0.0;
        e++;
      }
  lu_solve(A, b, 16);
  p.c = b;
  p.built = true;
}

void Spline3DO::init(double x0, double y0, double z0, const double v[2][2][2], const double dvdx[2][2][2],
                     const double dvdy[2][2][2], const double dvdz[2][2][2]) {
  p.setup(3, 3);
  p.x0[0] = x0;
  p.x0[1] = y0;
  p.x0[2] = z0;
  std::vector<double> A(64 * 64), b(64);
  int e = 0;
  // value, d/dx, d/dy, d/dz, then zero d2/dxdy, d2/dxdz, d2/dydz, d3/dxdydz
  const int ords[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                          {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (int q = 0; q < 8; q++)
    for (int i1 = 0; i1 < 2; i1++)
      for (int i2 = 0; i2 < 2; i2++)
        for (int i3 = 0; i3 < 2; i3++) {
          double xl[3] = {(double)i1, (double)i2, (double)i3};
          design_row(p, xl, ords[q], &A[e * 64]);
          double rhs = 0.0;
          if (q == 0) rhs = v[i1][i2][i3];
          if (q == 1) rhs = dvdx[i1][i2][i3];
          if (q == 2) rhs = dvdy[i1][i2][i3];
          if (q == 3) rhs = dvdz[i1][i2][i3];
          b[e] = rhs;
          e++;
        }
  lu_solve(A, b, 64);
  p.c = b;
  p.built = true;
}

void FuncGO::init(const double* x, const double* y, const double* dy, const double* d2y, int npts) {
  count = npts - 1;
  for (int i = 0; i < count; i++)
    sp[i].init(x[i], y[i], dy[i], d2y[i], x[i + 1], y[i + 1], dy[i + 1], d2y[i + 1]);
}

int FuncGO::cell(double c) const {  // AIREBO/cubici.hpp:82-100
  ORC_REQUIRE(c >= -1.0 && c <= 1.0);
  int s;
  for (s = 0; s < count; s++)
    if (c <= sp[s].xb) break;
  if (s >= count) s = count - 1;
  return s;
}

void Func2O::init(const double node[5][5], const double dx[5][5], const double dy[5][5], bool ia) {
  use_intarg = ia;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double z[2][2], zx[2][2], zy[2][2];
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          z[a][b] = node[i + a][j + b];
          zx[a][b] = dx[i + a][j + b];
          zy[a][b] = dy[i + a][j + b];
        }
      sp[i][j].init(i, j, z, zx, zy);
    }
  // fill_intarg, AIREBO/cubici.cxx:100-110
  for (int h = 0; h < 4; h++)
    for (int c = 0; c < 4; c++) {
      double x[2] = {(double)h, (double)c};
      intarg[h][c] = sp[h][c].p.eval(x);
      dH_intarg[h][c] = sp[h][c].p.eval(x, 0);
      dC_intarg[h][c] = sp[h][c].p.eval(x, 1);
    }
}

// AIREBO/cubici.hpp:153-222 (and Brenner/cubici.hpp:85-130 without intarg)
double Func2O::eval(double h, double c, int which) const {
  ORC_REQUIRE(h >= 0 && c >= 0);
  if (h >= 4 || c >= 4) return 0.0;
  int hi = int(h), ci = int(c);
  if (use_intarg && double(hi) == h && double(ci) == c)
    return which == 0 ? intarg[hi][ci] : which == 1 ? dH_intarg[hi][ci] : dC_intarg[hi][ci];
  double x[2] = {h, c};
  return sp[hi][ci].p.eval(x, which - 1);
}

void Func3O::init(int ni_, int nj_, int nk_, int built_k_, int kclamp_, bool ia,
                  const std::vector<double>& v, const std::vector<double>& vi,
                  const std::vector<double>& vj, const std::vector<double>& vk) {
  ni = ni_;
  nj = nj_;
  nk = nk_;
  built_k = built_k_;
  kclamp = kclamp_;
  use_intarg = ia;
  int nck = kclamp + 1;
  sp.assign((ni - 1) * (nj - 1) * nck, Spline3DO());
  for (auto& s : sp) s.p.setup(3, 3);
  auto N = [&](const std::vector<double>& a, int i, int j, int k) { return a[(i * nj + j) * nk + k]; };
  for (int i = 0; i < ni - 1; i++)
    for (int j = 0; j < nj - 1; j++)
      for (int k = 0; k < built_k; k++) {
        double c[2][2][2], ci[2][2][2], cj[2][2][2], ck[2][2][2];
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) {
              c[a][b][d] = N(v, i + a, j + b, k + d);
              ci[a][b][d] = N(vi, i + a, j + b, k + d);
              cj[a][b][d] = N(vj, i + a, j + b, k + d);
              ck[a][b][d] = N(vk, i + a, j + b, k + d);
            }
        sp[(i * (nj - 1) + j) * nck + k].init(i, j, k, c, ci, cj, ck);
      }
  int nint = (ni - 1) * (nj - 1) * nck;
  iav.assign(nint, 0.0);
  dia.assign(nint, 0.0);
  dja.assign(nint, 0.0);
  dka.assign(nint, 0.0);
  // fill_intarg, AIREBO/cubici.cxx:352-365 (cells beyond built_k are default: 0)
  for (int i = 0; i < ni - 1; i++)
    for (int j = 0; j < nj - 1; j++)
      for (int k = 0; k < nck; k++) {
        int id = (i * (nj - 1) + j) * nck + k;
        double x[3] = {(double)i, (double)j, (double)k};
        iav[id] = sp[id].p.eval(x);
        dia[id] = sp[id].p.eval(x, 0);
        dja[id] = sp[id].p.eval(x, 1);
        dka[id] = sp[id].p.eval(x, 2);
      }
}

// AIREBO/cubici.hpp:224-369 (Func_pi_rc, Func_Tij) and Brenner/cubici.hpp:132-175 (FuncF)
double Func3O::eval(double i, double j, double k, int which) const {
  ORC_REQUIRE(i >= 0 && j >= 0 && k >= 1);
  if (i >= ni - 1 || j >= nj - 1) return 0.0;
  if (k > kclamp) k = kclamp;
  int ii = int(i), jj = int(j), kk = int(k);
  int nck = kclamp + 1;
  int id = (ii * (nj - 1) + jj) * nck + kk;
  if (use_intarg && double(ii) == i && double(jj) == j && double(kk) == k) {
    switch (which) {
      case 0: return iav[id];
      case 1: return dia[id];
      case 2: return dja[id];
      default: return dka[id];
    }
  }
  double x[3] = {i, j, k};
  return sp[id].p.eval(x, which - 1);
}

}  // namespace orc
