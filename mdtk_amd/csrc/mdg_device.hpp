// Device-side types, parameter tables and inline math shared by the kernels.
//
// Layout: atoms are SoA in HBM (x,y,z,vx,... as separate double arrays),
// positions in cm (the reference's CGS units, mdtk/consts.cxx:28-38).  Atom
// type index: 0 = C, 1 = H, 2 = ion (Ar/Xe), 3 = other.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdg {

constexpr double kAo = 1e-8;
constexpr double kEV = 1.6021E-012;
constexpr double kAmu = 1.6605655e-24;
constexpr double kECharge = 4.8e-10;
constexpr double kNoPbc = 10000.0 * kAo;
constexpr double kPi = 3.14159265358979323846;

enum : int { T_C = 0, T_H = 1, T_ION = 2, T_OTHER = 3 };

// Error bits raised by kernels (checked by the host after a call).
enum : unsigned {
  ERR_SHORT_NL_OVERFLOW = 1u,
  ERR_LONG_NL_OVERFLOW = 2u,
  ERR_BOND_OVERFLOW = 4u,
  ERR_COS_RANGE = 8u,
  ERR_NEG_COORD = 16u,
  ERR_QUEUE_OVERFLOW = 32u,
  ERR_SPLINE_DOMAIN = 64u,
};

// Quintic Hermite pieces of G(cos theta), AIREBO/cubici.hpp:39-80.  Each piece
// is stored as local monomials p(t) = sum c_n t^n on [xa, xb], t = (x-xa)/h.
struct GSpline {
  int count;
  double xa[4], xb[4], h[4];
  double c[4][6];
};

// REBO parameters, mdtk/potentials/manybody/AIREBO/REBO.cxx:856-941.
struct ReboParams {
  double Q[2][2], alpha[2][2], A[2][2], B[3][2][2], beta[3][2][2], rho[2][2];
  double R[2][2][2], Rp[2][2][2], R2sq[2][2];
  double Ntot[2];
  GSpline gC1, gC2, gH;
};

// AIREBO-LJ parameters, AIREBO_LJ.cxx:219-287.
struct LJParams {
  double sigma[2][2], zeta[2][2], RLJ[2][2][2], R[2][2][2], R2sq[2][2], b[2][2][2];
};

// Brenner parameters, Brenner.cxx:497-611.
struct BrennerParams {
  double Re[2][2], De[2][2], beta[2][2], S[2][2], delta[2][2], R[2][2][2], R2sq[2][2], alpha[2][2][2];
  double ao, co, do_, G_HH;
};

// Node data of the bicubic and tricubic Hermite tables: value and first
// partial derivatives at integer nodes, cross derivatives zero
// (mdtk/Spline2D.cxx:35-125, mdtk/Spline3D.cxx:34-163).
struct Tables {
  ReboParams rebo;
  LJParams lj;
  BrennerParams bren;
  double zetaT[2][2];                 // torsion zeta, AIREBO_ETors.cxx:124-133
  double P[2][5][5][3];               // P_CC, P_CH: (v, d/dNH, d/dNC)
  double T3[4][5][5][16][4];          // pi_rc CC, CH, HH, T_ij: (v, di, dj, dk)
  double BH[2][5][5][3];              // Brenner H_CC, H_CH
  double BF[5][5][6][4];              // Brenner F (k nodes 0..5)
};

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double a, double b, double c) { return D3{a, b, c}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator-(D3 a) { return d3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ D3 operator/(D3 a, double s) { return d3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double norm2(D3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// Minimum-image component for a system whose atoms share one box
// (mdtk/Atom.hpp:115-140 with a1.PBC == a2.PBC): one +-L shift when |r| > L/2.
__device__ __forceinline__ double mic(double r, double L) {
  if (L != kNoPbc && fabs(r) > L * 0.5) r += (r > 0) ? (-L) : (L);
  return r;
}

struct Box {
  double L[3];
};

__device__ __forceinline__ D3 depos(const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, int a, int b, const Box& bx) {
  return d3(mic(x[a] - x[b], bx.L[0]), mic(y[a] - y[b], bx.L[1]), mic(z[a] - z[b], bx.L[2]));
}

// Cosine cutoff of AtomsPair (AtomsPair.hpp:142-154): f=1 below R1, cosine
// switch to 0 at R2, with df/dr.
__device__ __forceinline__ void cutoff(double r, double R1, double R2, double& f, double& df) {
  f = 0.0;
  df = 0.0;
  if (r < R1) {
    f = 1.0;
  } else if (r < R2) {
    double a = kPi * (r - R1) / (R2 - R1);
    f = (1.0 + cos(a)) / 2.0;
    df = (-(kPi / (R2 - R1)) * sin(a)) / 2.0;
  }
}

// REBO::Sprime / dSprime (REBO.cxx:67-106).
__device__ __forceinline__ double sprime(double a, double mn, double mx) {
  if (a < mn) return 1.0;
  if (a > mx) return 0.0;
  return (1.0 + cos(kPi * (a - mn) / (mx - mn))) / 2.0;
}
__device__ __forceinline__ double dsprime(double a, double mn, double mx) {
  if (a < mn || a > mx) return 0.0;
  return (-(kPi / (mx - mn)) * sin(kPi * (a - mn) / (mx - mn))) / 2.0;
}

// G(cos): first piece whose right end is >= c (cubici.hpp:82-100), Horner in t.
__device__ __forceinline__ void gspline(const GSpline& g, double c, double& v, double& d) {
  int s = 0;
  while (s < g.count - 1 && !(c <= g.xb[s])) s++;
  double t = (c - g.xa[s]) / g.h[s];
  const double* k = g.c[s];
  v = ((((k[5] * t + k[4]) * t + k[3]) * t + k[2]) * t + k[1]) * t + k[0];
  double dt = (((5.0 * k[5] * t + 4.0 * k[4]) * t + 3.0 * k[3]) * t + 2.0 * k[2]) * t + k[1];
  d = dt / g.h[s];
}

// Cubic Hermite basis on [0,1]: value basis h0 (node 0), h1 (node 1); slope
// basis s0, s1; and their derivatives.
__device__ __forceinline__ void hermite(double t, double h[2], double s[2], double dh[2], double ds[2]) {
  double t2 = t * t, t3 = t2 * t;
  h[0] = 2.0 * t3 - 3.0 * t2 + 1.0;
  h[1] = 3.0 * t2 - 2.0 * t3;
  s[0] = t3 - 2.0 * t2 + t;
  s[1] = t3 - t2;
  dh[0] = 6.0 * t2 - 6.0 * t;
  dh[1] = 6.0 * t - 6.0 * t2;
  ds[0] = 3.0 * t2 - 4.0 * t + 1.0;
  ds[1] = 3.0 * t2 - 2.0 * t;
}

// Bicubic P/H tables (cubici.hpp:153-222): zero for h>=4 or c>=4; node data
// stride (v, d/dh, d/dc).  out = {value, d/dh, d/dc}.
__device__ __forceinline__ void bicubic(const double (*nd)[5][3], double hh, double cc, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  if (hh >= 4.0 || cc >= 4.0) return;
  int i = int(hh), j = int(cc);
  double tx = hh - i, ty = cc - j;
  double hx[2], sx[2], dhx[2], dsx[2], hy[2], sy[2], dhy[2], dsy[2];
  hermite(tx, hx, sx, dhx, dsx);
  hermite(ty, hy, sy, dhy, dsy);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const double* q = nd[i + a][j + b];
      out[0] += q[0] * hx[a] * hy[b] + q[1] * sx[a] * hy[b] + q[2] * hx[a] * sy[b];
      out[1] += q[0] * dhx[a] * hy[b] + q[1] * dsx[a] * hy[b] + q[2] * dhx[a] * sy[b];
      out[2] += q[0] * hx[a] * dhy[b] + q[1] * sx[a] * dhy[b] + q[2] * hx[a] * dsy[b];
    }
}

// Tricubic tables.  nk = number of k nodes; kclamp = the reference's clamp of
// the third argument; kbuilt = number of built k cells (cells at and beyond it
// are default-constructed splines whose value is 0, cubici.cxx:343-345).
// out = {value, d/di, d/dj, d/dk}.
template <int NK>
__device__ __forceinline__ void tricubic(const double (*nd)[5][NK][4], double fi, double fj, double fk, int kclamp,
                                         int kbuilt, double out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (fi >= 4.0 || fj >= 4.0) return;
  if (fk > kclamp) fk = kclamp;
  int i = int(fi), j = int(fj), k = int(fk);
  if (k >= kbuilt) return;
  double tx = fi - i, ty = fj - j, tz = fk - k;
  double hx[2], sx[2], dhx[2], dsx[2], hy[2], sy[2], dhy[2], dsy[2], hz[2], sz[2], dhz[2], dsz[2];
  hermite(tx, hx, sx, dhx, dsx);
  hermite(ty, hy, sy, dhy, dsy);
  hermite(tz, hz, sz, dhz, dsz);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int c = 0; c < 2; c++) {
        const double* q = nd[i + a][j + b][k + c];
        double hyz = hy[b] * hz[c];
        out[0] += q[0] * hx[a] * hyz + q[1] * sx[a] * hyz + q[2] * hx[a] * sy[b] * hz[c] + q[3] * hx[a] * hy[b] * sz[c];
        out[1] += q[0] * dhx[a] * hyz + q[1] * dsx[a] * hyz + q[2] * dhx[a] * sy[b] * hz[c] +
                  q[3] * dhx[a] * hy[b] * sz[c];
        out[2] += q[0] * hx[a] * dhy[b] * hz[c] + q[1] * sx[a] * dhy[b] * hz[c] + q[2] * hx[a] * dsy[b] * hz[c] +
                  q[3] * hx[a] * dhy[b] * sz[c];
        out[3] += q[0] * hx[a] * hy[b] * dhz[c] + q[1] * sx[a] * hy[b] * dhz[c] + q[2] * hx[a] * sy[b] * dhz[c] +
                  q[3] * hx[a] * hy[b] * dsz[c];
      }
}

__device__ __forceinline__ void atomic_add3(double* g, int a, D3 v) {
  unsafeAtomicAdd(g + 3 * a + 0, v.x);
  unsafeAtomicAdd(g + 3 * a + 1, v.y);
  unsafeAtomicAdd(g + 3 * a + 2, v.z);
}

}  // namespace mdg
