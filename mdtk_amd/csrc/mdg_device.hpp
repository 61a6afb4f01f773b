// Device-side types, parameter tables and inline math shared by the kernels.
//
// Layout: atoms are SoA in HBM (x,y,z,vx,... as separate double arrays),
// positions in cm (the reference's CGS units, mdtk/consts.cxx:28-38).  Atom
// type index: 0 = C, 1 = H, 2 = ion (Ar/Xe), 3 = other.
//
// Force kernels are templated on the geometry type R.  R = double works in the
// reference's CGS units (cm, erg) with the reference's operation order; R =
// float works in Angstrom and eV on displacements taken in FP64 and rounded
// once to FP32 (a per-pair local frame), while coordination sums, D/Nconj
// sums, energies and force accumulation stay FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace mdg {

constexpr double kAo = 1e-8;
constexpr double kEV = 1.6021E-012;
constexpr double kAmu = 1.6605655e-24;
constexpr double kECharge = 4.8e-10;
constexpr double kNoPbc = 10000.0 * kAo;
constexpr double kPi = 3.14159265358979323846;
constexpr int kNumTheta = 33;  // REBO pair (3 x 9) and LJ (3 x 2) parameters per system

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
  ERR_EAM_ROW_OVERFLOW = 128u,
  ERR_PATH_OVERFLOW = 256u,   // an atom's LJ pairs or C_ij paths beyond their per-atom capacity
  ERR_TYPE_RANGE = 512u,      // a packed position carries a type outside the tables
};
// Errors that a global resource (the b* queue, the bond array) raises; every
// other bit belongs to one system.
constexpr unsigned kGlobalErrors = ERR_QUEUE_OVERFLOW;

// Device error words: w[0] collects errors of global resources, w[1 + s]
// those of system s (checked by the host after each call).
__device__ __forceinline__ unsigned* sys_err(unsigned* w, int s) { return w + 1 + s; }
__device__ __forceinline__ void raise(unsigned* e, unsigned bit) { atomicOr(e, bit); }

// Quintic Hermite pieces of G(cos theta), AIREBO/cubici.hpp:39-80.  Each piece
// is stored as local monomials p(t) = sum c_n t^n on [xa, xb], t = (x-xa)/h.
template <class R>
struct GSplineT {
  int count;
  R xa[4], xb[4], h[4];
  R c[4][6];
};

// REBO parameters, mdtk/potentials/manybody/AIREBO/REBO.cxx:856-941.
template <class F>
struct ReboParamsT {
  F Q[2][2], alpha[2][2], A[2][2], B[3][2][2], beta[3][2][2], rho[2][2];
  F R[2][2][2], Rp[2][2][2], R2sq[2][2];
  F Ntot[2];
  F expscale;  // 4/Angstrom of ExpTerm (REBO.cxx:221) in the table's length unit
  GSplineT<F> gC1, gC2, gH;
};

// AIREBO-LJ parameters, AIREBO_LJ.cxx:219-287.
template <class F>
struct LJParamsT {
  F sigma[2][2], zeta[2][2], RLJ[2][2][2], R[2][2][2], R2sq[2][2], b[2][2][2];
};

// Brenner parameters, Brenner.cxx:497-611.
template <class F>
struct BrennerParamsT {
  F Re[2][2], De[2][2], beta[2][2], S[2][2], delta[2][2], R[2][2][2], R2sq[2][2], alpha[2][2][2];
  F ao, co, do_, G_HH;
};

// Node data of the bicubic and tricubic Hermite tables: value and first
// partial derivatives at integer nodes, cross derivatives zero
// (mdtk/Spline2D.cxx:35-125, mdtk/Spline3D.cxx:34-163).
template <class R>
struct TablesT {
  double lunit;  // cm per length unit of this table
  double eunit;  // erg per energy unit of this table
  ReboParamsT<R> rebo;
  LJParamsT<R> lj;
  BrennerParamsT<R> bren;
  R zetaT[2][2];               // torsion zeta, AIREBO_ETors.cxx:124-133
  R P[2][5][5][3];             // P_CC, P_CH: (v, d/dNH, d/dNC)
  R T3[4][5][5][16][4];        // pi_rc CC, CH, HH, T_ij: (v, di, dj, dk)
  R BH[2][5][5][3];            // Brenner H_CC, H_CH
  R BF[5][5][6][4];            // Brenner F (k nodes 0..5)
  // Upper bound of pi_rc over its whole domain for each (type i, type j)
  // orientation (the bicubic/tricubic interpolant can overshoot its nodes
  // only through node slopes; see rc_upper_bound), plus a margin for the
  // FP32 evaluation.  Used to prove b* < b_min without evaluating pi_rc.
  double rcmax[2][2];
  // Lower bounds over the whole domain, minus a margin for the FP32 form's
  // evaluation: glb[0] of G for a C centre (both quintics G_C1, G_C2; their
  // S'(N) blend lies between them), glb[1] of G_H; plb[t] of P_CC (t = C) and
  // P_CH (t = H), the out-of-table zero included.  With them the coordination
  // pass bounds, for every atom a and partner type t, D_a^-1/2 from above for
  // any partner direction (pmax), so that b* < b_min is proven per pair
  // without its sigma-pi sums.
  double glb[2], plb[2];
};
// FLJ (pairwise/FLJ.cxx): element index e2i (Cu 0, H 1, C 2, Au 3), LB-mixed
// sigma/epsilon, the ZBL-to-LJ cubic bridge ends x[0..1] with the values and
// slopes of each element pair, and the cosine cutoff R0..R1 (cm, erg).
struct FLJTable {
  double sigma[4][4], eps[4][4];
  double x[2];
  double v[4][4][2], d[4][4][2];
  double R0, R1;
};
void build_flj_table(FLJTable& t);

// Ackland EAM (manybody/Ackland/Ackland.{hpp,cxx}) for Cu, Ag, Au, indexed
// Cu = 0, Ag = 1, Au = 2.  Pair term V: knots rk/ak (k = 1..6 like, 1..3
// unlike; V = 0 from rk[1] on), ZBL below x[0] with the screening length
// AS, the cubic through both ends' values and slopes on [x[0], x[1]).
// Density phi_e: knots Rk/Ak (k = 1, 2) of the element; unlike pairs take
// the geometric mean of the two elements' phi.
struct EAMTable {
  double rk[3][3][7], ak[3][3][7];
  double Rk[3][2], Ak[3][2];
  double AS[3][3];
  double Z[3];  // nuclear charges (atomic number x e)
  double x[2], v[3][3][2], d[3][3][2];
  double rmax;  // no pair term or density beyond this distance
};
void build_eam_table(EAMTable& t);

// TightBinding (manybody/TightBinding/TightBinding.cxx), Cu only: density
// pair function f(r) exp(-beta r); pair term Born-Mayer A exp(-B r) below
// x[0], the cubic through both ends' values and slopes to x[1], f(r) Phi0
// exp(-alpha r) beyond; f the cosine cutoff R0..R1; embedding -c sqrt(rho).
struct TBTable {
  double alpha, beta, c, Phi0, R0, R1, A, B;
  double x[2], v[2], d[2];
};
void build_tb_table(TBTable& t);

// FBM (pairwise/FBM.cxx): A exp(-B r) times the cosine cutoff R0..R1 for
// Cu-Cu and Cu-Ar (index Cu = 0, Ar = 1; Ar-Ar is not a pair of it).
struct FBMTable {
  double A[2][2], B[2][2];
  double R0, R1;
};
void build_fbm_table(FBMTable& t);

using Tables = TablesT<double>;
using ReboParams = ReboParamsT<double>;
using LJParams = LJParamsT<double>;
using BrennerParams = BrennerParamsT<double>;
using GSpline = GSplineT<double>;

template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
__device__ __forceinline__ V3<R> v3(R a, R b, R c) {
  return V3<R>{a, b, c};
}
template <class R>
__device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) {
  return v3<R>(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class R>
__device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) {
  return v3<R>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class R>
__device__ __forceinline__ V3<R> operator-(V3<R> a) {
  return v3<R>(-a.x, -a.y, -a.z);
}
template <class R>
__device__ __forceinline__ V3<R> operator*(V3<R> a, R s) {
  return v3<R>(a.x * s, a.y * s, a.z * s);
}
template <class R>
__device__ __forceinline__ V3<R> operator/(V3<R> a, R s) {
  return v3<R>(a.x / s, a.y / s, a.z / s);
}
template <class R>
__device__ __forceinline__ R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
__device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b) {
  return v3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R>
__device__ __forceinline__ R norm2(V3<R> a) {
  return a.x * a.x + a.y * a.y + a.z * a.z;
}

using D3 = V3<double>;
__device__ __forceinline__ D3 d3(double a, double b, double c) { return D3{a, b, c}; }

// Minimum-image component for a system whose atoms share one box
// (mdtk/Atom.hpp:115-140 with a1.PBC == a2.PBC): one +-L shift when |r| > L/2.
__device__ __forceinline__ double mic(double r, double L) {
  if (L != kNoPbc && fabs(r) > L * 0.5) r += (r > 0) ? (-L) : (L);
  return r;
}

// Boxes of a batch of independent systems: L[3*s..3*s+2] is system s's box,
// sys[a] the system of atom a (all zero for a single system).
// Atoms carry their own PBC vectors in the reference (Atom::PBC).  A system
// whose atoms do not all share its box is "mixed": on each periodic axis of
// the box some of its atoms are open (NO_PBC), as a state file's per-atom
// enabled flags allow; open_axes[a] bit d marks them.  Only the list build,
// the packed copy and the PBC wrap look at this (see packed_coord below);
// apbc == null when no system is mixed.
struct Box {
  const double* L;
  const int* sys;
  const unsigned char* apbc;       // per system
  const unsigned char* open_axes;  // per atom
  const int *pcx, *pcy, *pcz;  // Atom::PBC_count, for the unwrap of mixed pairs
};
__device__ __forceinline__ const double* box_of(const Box& bx, int a) { return bx.L + 3 * bx.sys[a]; }

// One system's periodicity as its atom a sees it.
struct Frame {
  const double* L;
};
__device__ __forceinline__ Frame frame_of(const Box& bx, int a) { return Frame{bx.L + 3 * bx.sys[a]}; }

// The PBC value of atom a's axis d (its own Atom::PBC).
__device__ __forceinline__ double pbc_of(const Box& bx, const double* L, int a, int d) {
  return (bx.apbc && bx.apbc[bx.sys[a]] && ((bx.open_axes[a] >> d) & 1)) ? kNoPbc : L[d];
}

// Atom a's own PBC vector.
struct Pbc3 {
  double v[3];
};
__device__ __forceinline__ Pbc3 own_pbc(const Box& bx, int a) {
  const double* L = box_of(bx, a);
  Pbc3 o;
#pragma unroll
  for (int d = 0; d < 3; d++) o.v[d] = pbc_of(bx, L, a, d);
  return o;
}

// One axis of depos for atoms with PBC values La, Lb and image counters ca,
// cb (mdtk/Atom.hpp:121-133): an open atom sees a periodic one at its
// unwrapped coordinate; two atoms of one periodicity take the minimum image.
__device__ __forceinline__ double depos_axis(double r, double La, double Lb, const int* pc, int a, int b) {
  if (La == kNoPbc && Lb != kNoPbc) r -= Lb * pc[b];
  if (Lb == kNoPbc && La != kNoPbc) r += La * pc[a];
  if (La == Lb && La != kNoPbc && fabs(r) > La * 0.5) r += (r > 0) ? (-La) : (La);
  return r;
}

// depos with both atoms' own PBC vectors from their true coordinates
// (A.x/y/z, not the packed copy): the list build of mixed systems.
__device__ __forceinline__ D3 depos_atoms(const Box& bx, const double* x, const double* y, const double* z, int a,
                                          int b) {
  const double* L = box_of(bx, a);
  return d3(depos_axis(x[a] - x[b], pbc_of(bx, L, a, 0), pbc_of(bx, L, b, 0), bx.pcx, a, b),
            depos_axis(y[a] - y[b], pbc_of(bx, L, a, 1), pbc_of(bx, L, b, 1), bx.pcy, a, b),
            depos_axis(z[a] - z[b], pbc_of(bx, L, a, 2), pbc_of(bx, L, b, 2), bx.pcz, a, b));
}

// The packed copy holds every coordinate reduced into [0, L) on each
// periodic axis of the system's box.  For the atoms whose own PBC is the box
// that is their wrapped coordinate; an atom of a mixed system that is open
// on such an axis keeps its true coordinate in A.x and its reduced one here.
// The force kernels then take plain minimum images, and for every pair the
// mixed list build lists, that is the reference's depos: the two differ by a
// multiple of L, and a listed pair stays within 2 (1 + 0.44) Rc of its
// partner, so the minimum image cannot pick another copy within a cutoff
// when L >= 2 (1 + 0.44) Rc (mdg_set_atom_pbc requires it of mixed systems).
__device__ __forceinline__ double packed_coord(double x, double L) {
  if (L == kNoPbc || (x >= 0.0 && x < L)) return x;
  double q = x - L * floor(x / L);
  return (q >= L) ? q - L : q;
}

// x_a - x_b (mdtk/Atom.hpp:115-140) for atoms a, b of the system of frame f,
// from their packed positions: the minimum image.
__device__ __forceinline__ D3 disp(const Box&, const Frame& f, int, int, double4 pa, double4 pb) {
  return d3(mic(pa.x - pb.x, f.L[0]), mic(pa.y - pb.y, f.L[1]), mic(pa.z - pb.z, f.L[2]));
}

// Positions are gathered from a packed copy p[a] = (x, y, z, type): one
// 32-byte load per atom instead of three scattered ones.
__device__ __forceinline__ D3 depos(const double4* __restrict__ p, int a, int b, const Box& bx) {
  return disp(bx, frame_of(bx, a), a, b, p[a], p[b]);
}

// A displacement taken in FP64 cm, expressed in the kernel's unit and type:
// identity for R = double, one rounding to FP32 Angstrom for R = float.
template <class R>
__device__ __forceinline__ V3<R> to_local(D3 d) {
  if constexpr (std::is_same<R, double>::value) {
    return d;
  } else {
    return v3<R>(R(d.x * 1e8), R(d.y * 1e8), R(d.z * 1e8));
  }
}

// cos(pi t) and sin(pi t) for t in [0, 1], FP32, for the mixed form's
// switching functions.  With x = pi (t - 1/2), |x| <= pi/2:
// cos(pi t) = -sin x and sin(pi t) = cos x, from the Taylor series of sin x
// through x^13 and of cos x through x^14 by Horner's rule in x^2 (the first
// omitted terms are below 1e-9, under an FP32 ulp of the results' scale).
// The library's cosf/sinf reduce arbitrary arguments and cost about ten
// times as many instructions in every wave that reaches a switch.
__host__ __device__ __forceinline__ void cospi_sinpi_unit(float t, float& c, float& s) {
  const float x = float(kPi) * (t - 0.5f);
  const float x2 = x * x;
  float ps = 1.6059043836821615e-10f;  // 1/13!
  ps = ps * x2 - 2.5052108385441719e-08f;
  ps = ps * x2 + 2.7557319223985891e-06f;
  ps = ps * x2 - 1.9841269841269841e-04f;
  ps = ps * x2 + 8.3333333333333333e-03f;
  ps = ps * x2 - 1.6666666666666667e-01f;
  ps = ps * x2 + 1.0f;
  float pc = -1.1470745597729725e-11f;  // -1/14!
  pc = pc * x2 + 2.0876756987868099e-09f;
  pc = pc * x2 - 2.7557319223985891e-07f;
  pc = pc * x2 + 2.4801587301587302e-05f;
  pc = pc * x2 - 1.3888888888888889e-03f;
  pc = pc * x2 + 4.1666666666666667e-02f;
  pc = pc * x2 - 0.5f;
  pc = pc * x2 + 1.0f;
  // kept inside [-1, 1] and [0, 1] as the functions are (a switch f = (1 + c)/2
  // never goes below zero)
  c = fminf(fmaxf(-(x * ps), -1.0f), 1.0f);
  s = fmaxf(pc, 0.0f);
}
// The same in FP64 (sin through x^21, cos through x^22; first omitted terms
// below 2e-18), for the mixed form's FP64 pair terms.
__host__ __device__ __forceinline__ void cospi_sinpi_unit(double t, double& c, double& s) {
  const double x = kPi * (t - 0.5);
  const double x2 = x * x;
  double ps = 1.95729410633912626e-20;  // 1/21!
  ps = ps * x2 - 8.22063524662432950e-18;
  ps = ps * x2 + 2.81145725434552060e-15;
  ps = ps * x2 - 7.64716373181981641e-13;
  ps = ps * x2 + 1.60590438368216133e-10;
  ps = ps * x2 - 2.50521083854417202e-08;
  ps = ps * x2 + 2.75573192239858925e-06;
  ps = ps * x2 - 1.98412698412698413e-04;
  ps = ps * x2 + 8.33333333333333322e-03;
  ps = ps * x2 - 1.66666666666666657e-01;
  ps = ps * x2 + 1.0;
  double pc = 8.89679139245057408e-22;  // 1/22!
  pc = pc * x2 - 4.11031762331216484e-19;
  pc = pc * x2 + 1.56192069685862253e-16;
  pc = pc * x2 - 4.77947733238738525e-14;
  pc = pc * x2 + 1.14707455977297245e-11;
  pc = pc * x2 - 2.08767569878681002e-09;
  pc = pc * x2 + 2.75573192239858883e-07;
  pc = pc * x2 - 2.48015873015873016e-05;
  pc = pc * x2 + 1.38888888888888894e-03;
  pc = pc * x2 - 4.16666666666666644e-02;
  pc = pc * x2 + 0.5;
  pc = 1.0 - pc * x2;
  c = fmin(fmax(-(x * ps), -1.0), 1.0);
  s = fmax(pc, 0.0);
}

// Cosine cutoff of AtomsPair (AtomsPair.hpp:142-154): f=1 below R1, cosine
// switch to 0 at R2, with df/dr.  POLY: the switch from cospi_sinpi_unit
// (the mixed form; FP32 kernels, and their FP64 pair terms), else from the
// library's cos and sin as the reference computes it (the FP64 form).
template <class R, bool POLY = std::is_same<R, float>::value>
__device__ __forceinline__ void cutoff(R r, R R1, R R2, R& f, R& df) {
  f = R(0);
  df = R(0);
  if (r < R1) {
    f = R(1);
  } else if (r < R2) {
    if constexpr (POLY) {
      R c, sn;
      cospi_sinpi_unit((r - R1) / (R2 - R1), c, sn);
      f = (R(1) + c) / R(2);
      df = (-(R(kPi) / (R2 - R1)) * sn) / R(2);
    } else {
      R a = R(kPi) * (r - R1) / (R2 - R1);
      f = (R(1) + cos(a)) / R(2);
      df = (-(R(kPi) / (R2 - R1)) * sin(a)) / R(2);
    }
  }
}

// REBO::Sprime / dSprime (REBO.cxx:67-106).
template <class R>
__device__ __forceinline__ R sprime(R a, R mn, R mx) {
  if (a < mn) return R(1);
  if (a > mx) return R(0);
  if constexpr (std::is_same<R, float>::value) {
    float c, sn;
    cospi_sinpi_unit((a - mn) / (mx - mn), c, sn);
    return (1.0f + c) / 2.0f;
  } else {
    return (R(1) + cos(R(kPi) * (a - mn) / (mx - mn))) / R(2);
  }
}
template <class R>
__device__ __forceinline__ R dsprime(R a, R mn, R mx) {
  if (a < mn || a > mx) return R(0);
  if constexpr (std::is_same<R, float>::value) {
    float c, sn;
    cospi_sinpi_unit((a - mn) / (mx - mn), c, sn);
    return (-(float(kPi) / (mx - mn)) * sn) / 2.0f;
  } else {
    return (-(R(kPi) / (mx - mn)) * sin(R(kPi) * (a - mn) / (mx - mn))) / R(2);
  }
}

// G(cos): first piece whose right end is >= c (cubici.hpp:82-100), Horner in t.
template <class R>
__device__ __forceinline__ void gspline(const GSplineT<R>& g, R c, R& v, R& d) {
  int s = 0;
  while (s < g.count - 1 && !(c <= g.xb[s])) s++;
  R t = (c - g.xa[s]) / g.h[s];
  const R* k = g.c[s];
  v = ((((k[5] * t + k[4]) * t + k[3]) * t + k[2]) * t + k[1]) * t + k[0];
  R dt = (((R(5) * k[5] * t + R(4) * k[4]) * t + R(3) * k[3]) * t + R(2) * k[2]) * t + k[1];
  d = dt / g.h[s];
}

// Cubic Hermite basis on [0,1]: value basis h0 (node 0), h1 (node 1); slope
// basis s0, s1; and their derivatives.
template <class R>
__device__ __forceinline__ void hermite(R t, R h[2], R s[2], R dh[2], R ds[2]) {
  R t2 = t * t, t3 = t2 * t;
  h[0] = R(2) * t3 - R(3) * t2 + R(1);
  h[1] = R(3) * t2 - R(2) * t3;
  s[0] = t3 - R(2) * t2 + t;
  s[1] = t3 - t2;
  dh[0] = R(6) * t2 - R(6) * t;
  dh[1] = R(6) * t - R(6) * t2;
  ds[0] = R(3) * t2 - R(4) * t + R(1);
  ds[1] = R(3) * t2 - R(2) * t;
}

// Bicubic P/H tables (cubici.hpp:153-222): zero for h>=4 or c>=4; node data
// stride (v, d/dh, d/dc).  out = {value, d/dh, d/dc}.
template <class R>
__device__ __forceinline__ void bicubic(const R (*nd)[5][3], R hh, R cc, R out[3]) {
  out[0] = out[1] = out[2] = R(0);
  if (hh >= R(4) || cc >= R(4)) return;
  int i = int(hh), j = int(cc);
  R tx = hh - R(i), ty = cc - R(j);
  R hx[2], sx[2], dhx[2], dsx[2], hy[2], sy[2], dhy[2], dsy[2];
  hermite(tx, hx, sx, dhx, dsx);
  hermite(ty, hy, sy, dhy, dsy);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const R* q = nd[i + a][j + b];
      out[0] += q[0] * hx[a] * hy[b] + q[1] * sx[a] * hy[b] + q[2] * hx[a] * sy[b];
      out[1] += q[0] * dhx[a] * hy[b] + q[1] * dsx[a] * hy[b] + q[2] * dhx[a] * sy[b];
      out[2] += q[0] * hx[a] * dhy[b] + q[1] * sx[a] * dhy[b] + q[2] * hx[a] * dsy[b];
    }
}

// Tricubic tables.  kclamp = the reference's clamp of the third argument;
// kbuilt = number of built k cells (cells at and beyond it are
// default-constructed splines whose value is 0, cubici.cxx:343-345).
// out = {value, d/di, d/dj, d/dk}.
template <class R, int NK>
__device__ __forceinline__ void tricubic(const R (*nd)[5][NK][4], R fi, R fj, R fk, int kclamp, int kbuilt,
                                         R out[4]) {
  out[0] = out[1] = out[2] = out[3] = R(0);
  if (fi >= R(4) || fj >= R(4)) return;
  if (fk > R(kclamp)) fk = R(kclamp);
  int i = int(fi), j = int(fj), k = int(fk);
  if (k >= kbuilt) return;
  R tx = fi - R(i), ty = fj - R(j), tz = fk - R(k);
  R hx[2], sx[2], dhx[2], dsx[2], hy[2], sy[2], dhy[2], dsy[2], hz[2], sz[2], dhz[2], dsz[2];
  hermite(tx, hx, sx, dhx, dsx);
  hermite(ty, hy, sy, dhy, dsy);
  hermite(tz, hz, sz, dhz, dsz);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int c = 0; c < 2; c++) {
        const R* q = nd[i + a][j + b][k + c];
        R hyz = hy[b] * hz[c];
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

// Deterministic accumulation of the FP64 form.  Floating-point atomics add
// in whatever order the hardware takes them, so FP64 runs would not repeat
// bit for bit.  The FP64 form instead splits every contribution v exactly
// into hi * 2^-30 + lo * 2^-60 (hi = trunc(v 2^30); the remainder is exact by
// Sterbenz's lemma and lo rounds it once at 2^-60) and adds the two integers
// with integer atomics, whose sums do not depend on the order; the gradient
// (erg/cm) is read back as hi 2^-30 + lo 2^-60 once per force call: range
// 2^33, resolution 2^-60 (5e-15 eV/A).  Energies (erg) are scaled by 2^40
// first (resolution 8e-31 erg).
constexpr double kFixHi = 1073741824.0;           // 2^30
constexpr double kFixLo = 1152921504606846976.0;  // 2^60
constexpr double kFixEnergy = 1099511627776.0;    // 2^40
__device__ __forceinline__ void fix_split(double v, long long& hi, long long& lo) {
  const double h = trunc(v * kFixHi);
  hi = (long long)h;
  lo = llrint((v - h * (1.0 / kFixHi)) * kFixLo);
}
__device__ __forceinline__ double fix_join(long long hi, long long lo) {
  return double(hi) * (1.0 / kFixHi) + double(lo) * (1.0 / kFixLo);
}
__device__ __forceinline__ void fix_atomic(long long* p, double v) {
  if (v == 0.0) return;
  long long hi, lo;
  fix_split(v, hi, lo);
  if (hi) atomicAdd((unsigned long long*)p, (unsigned long long)hi);
  if (lo) atomicAdd((unsigned long long*)(p + 1), (unsigned long long)lo);
}
// fixed-point gradient: per atom and component the pair (hi, lo)
__device__ __forceinline__ void fix_add3(long long* g, int a, D3 v) {
  fix_atomic(g + 6 * a + 0, v.x);
  fix_atomic(g + 6 * a + 2, v.y);
  fix_atomic(g + 6 * a + 4, v.z);
}

// The gradient of kernels without a workgroup cache: FP64 atomics into grad,
// or the fixed-point pairs when fix is set (FP64 form).
struct GradSink {
  double* grad;
  long long* fix;
};
__device__ __forceinline__ void atomic_add3(const GradSink& g, int a, D3 v) {
  if (g.fix) fix_add3(g.fix, a, v);
  else atomic_add3(g.grad, a, v);
}

// Workgroup gradient cache: S slots of LDS keyed by atom index (linear
// probing from a multiplicative hash).  A scatter finds or claims its atom's
// slot and adds there with LDS atomics (FP64, or the fixed-point pairs when
// FIX); after 8 probes it adds into the global gradient instead.  flush()
// moves every claimed slot into the global gradient with one atomic per
// component, so the atoms that a workgroup's bonds or pairs share cost one
// global update per flush instead of one per contribution.  clear() and
// flush() are called by the whole workgroup with a barrier between them and
// any add().
#define MDG_LDS __attribute__((address_space(3)))
template <int S, bool FIX = false>
struct GradCache {
  static_assert((S & (S - 1)) == 0 && S >= 64, "S: power of two");
  using val_t = typename std::conditional<FIX, long long, double>::type;
  static constexpr int kSlots = S;
  static constexpr int kPer = FIX ? 6 : 3;  // values per slot
  static constexpr int kVals = kPer * S;
  MDG_LDS int* key;     // [S], -1 when free
  MDG_LDS val_t* val;   // [kPer S]
  MDG_LDS int* fill;    // slots claimed since the last clear
  double* grad;         // global gradient, n*3
  long long* gfix;      // fixed-point global gradient, n*6 (FIX)

  __device__ __forceinline__ GradCache(int* k, val_t* v, int* f, double* g, long long* gf)
      : key((MDG_LDS int*)k), val((MDG_LDS val_t*)v), fill((MDG_LDS int*)f), grad(g), gfix(gf) {}

  __device__ __forceinline__ void clear() const {
    for (int q = threadIdx.x; q < S; q += blockDim.x) {
      key[q] = -1;
#pragma unroll
      for (int c = 0; c < kPer; c++) val[kPer * q + c] = val_t(0);
    }
    if (threadIdx.x == 0) *fill = 0;
  }
  __device__ __forceinline__ void lds_add(int h, double x, double y, double z) const {
    if constexpr (FIX) {
      const double v[3] = {x, y, z};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (v[c] == 0.0) continue;
        long long hi, lo;
        fix_split(v[c], hi, lo);
        if (hi) __hip_atomic_fetch_add(&val[6 * h + 2 * c], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (lo) __hip_atomic_fetch_add(&val[6 * h + 2 * c + 1], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    } else {
      __hip_atomic_fetch_add(&val[3 * h], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&val[3 * h + 1], y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&val[3 * h + 2], z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __device__ __forceinline__ void global_add(int a, double x, double y, double z) const {
    if constexpr (FIX) fix_add3(gfix, a, d3(x, y, z));
    else atomic_add3(grad, a, d3(x, y, z));
  }
  __device__ __forceinline__ void add(int a, double x, double y, double z) const {
    constexpr int kBits = __builtin_ctz(S);
    int h = (int)(((unsigned)a * 2654435761u) >> (32 - kBits));
#pragma unroll 1
    for (int p = 0; p < 8; p++) {
      int k = __hip_atomic_load(&key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (k == -1) {
        int expect = -1;
        if (__hip_atomic_compare_exchange_strong(&key[h], &expect, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP)) {
          k = a;
          __hip_atomic_fetch_add(fill, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
          k = expect;
        }
      }
      if (k == a) {
        lds_add(h, x, y, z);
        return;
      }
      h = (h + 1) & (S - 1);
    }
    global_add(a, x, y, z);
  }
  __device__ __forceinline__ void flush() const {
    for (int q = threadIdx.x; q < S; q += blockDim.x) {
      int a = key[q];
      if (a < 0) continue;
      if constexpr (FIX) {
#pragma unroll
        for (int c = 0; c < 6; c++)
          if (val[6 * q + c])
            atomicAdd((unsigned long long*)(gfix + 6 * a + c), (unsigned long long)val[6 * q + c]);
      } else {
        atomic_add3(grad, a, d3(val[3 * q], val[3 * q + 1], val[3 * q + 2]));
      }
    }
  }
  // Called by the whole workgroup after a barrier: empties the cache when it
  // is more than half full (the decision reads one LDS value, the same for
  // every thread).
  __device__ __forceinline__ void drain_if_crowded() const {
    if (*fill > S / 2) {
      __syncthreads();
      flush();
      __syncthreads();
      clear();
    }
    __syncthreads();
  }
};

template <int S, bool FIX>
__device__ __forceinline__ void atomic_add3(const GradCache<S, FIX>& g, int a, D3 v) {
  g.add(a, v.x, v.y, v.z);
}

// The cache a kernel of geometry type R uses, in the LDS of a cache of S FP64
// slots: the FP64 form (R = double) accumulates in fixed point (half the
// slots, each twice as wide).
template <class R, int S>
using CacheOf = typename std::conditional<std::is_same<R, double>::value, GradCache<S / 2, true>,
                                          GradCache<S, false>>::type;

// Gradient scatter from a kernel working in unit/type R into the FP64 CGS
// gradient (global, or through a workgroup cache): gs = (erg per energy
// unit) / (cm per length unit).
template <class R, class G>
__device__ __forceinline__ void scatter3(const G& g, int a, V3<R> v, double gs) {
  if constexpr (std::is_same<R, double>::value) {
    atomic_add3(g, a, v);
  } else {
    atomic_add3(g, a, d3(double(v.x) * gs, double(v.y) * gs, double(v.z) * gs));
  }
}

// Virial of one kernel's terms, W_ab = sum over pair pushes of -r_a v_b,
// where a push adds +v to atom a and -v to atom b and r = x_a - x_b (the
// minimum-image vector the term itself uses): every force term here is a
// sum of such pushes along its own pair vectors, so W is the virial of the
// forces as computed, for any box.  Components xx, yy, zz, xy, xz, yz; the
// kernel's energy unit, FP64.  Accumulated per thread for the system of the
// current term and added into vir[6 s + c] (erg) when the system changes and
// at the end.  Virial<false> compiles to nothing.
template <bool ON>
struct Virial {
  __device__ __forceinline__ Virial(long long* = nullptr, double = 1.0) {}
  __device__ __forceinline__ void at(int) {}
  template <class R>
  __device__ __forceinline__ void push(V3<R>, V3<R>) {}
  template <class R>
  __device__ __forceinline__ void push_half(V3<R>, V3<R>) {}
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ void flush_wave() {}
  template <int LW>
  __device__ __forceinline__ void flush_group(int) {}
};

template <>
struct Virial<true> {
  double w[6] = {0, 0, 0, 0, 0, 0};
  int sys = -1;
  long long* out = nullptr;  // per system six fixed-point (hi, lo) pairs of the energy scale
  double scale = 1.0;
  __device__ __forceinline__ Virial(long long* o = nullptr, double sc = 1.0) : out(o), scale(sc) {}
  // Integer sums, so the virial of a force call does not depend on the order
  // in which waves add their parts (as the fixed-point gradient).
  __device__ __forceinline__ void add_out() {
    if (sys < 0) return;
#pragma unroll
    for (int c = 0; c < 6; c++)
      if (w[c] != 0.0) fix_atomic(out + 12 * sys + 2 * c, w[c] * scale * kFixEnergy);
#pragma unroll
    for (int c = 0; c < 6; c++) w[c] = 0.0;
  }
  __device__ __forceinline__ void at(int s) {
    if (s != sys) {
      add_out();
      sys = s;
    }
  }
  template <class R>
  __device__ __forceinline__ void push(V3<R> r, V3<R> v) {
    const double rx = double(r.x), ry = double(r.y), rz = double(r.z);
    const double vx = double(v.x), vy = double(v.y), vz = double(v.z);
    w[0] -= rx * vx;
    w[1] -= ry * vy;
    w[2] -= rz * vz;
    w[3] -= rx * vy;
    w[4] -= rx * vz;
    w[5] -= ry * vz;
  }
  // A pair visited from both of its atoms (gather kernels): each visit
  // pushes half.
  template <class R>
  __device__ __forceinline__ void push_half(V3<R> r, V3<R> v) {
    push<R>(r * R(0.5), v);
  }
  __device__ __forceinline__ void flush() { add_out(); }
  // A group of LW consecutive lanes working on one atom (same system): sum
  // over the group, lane sl == 0 adds it.
  template <int LW>
  __device__ __forceinline__ void flush_group(int sl) {
#pragma unroll
    for (int c = 0; c < 6; c++) {
      double v = w[c];
      for (int o = LW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LW);
      w[c] = (sl == 0) ? v : 0.0;
    }
    add_out();
  }
  // Whole-wave flush (every lane of the wave must call it): one set of
  // atomics per wave when the lanes that hold terms share one system.
  __device__ __forceinline__ void flush_wave() {
    int s0 = sys;
    for (int o = 32; o > 0; o >>= 1) s0 = max(s0, __shfl_xor(s0, o, 64));
    if (s0 >= 0 && __all(sys == s0 || sys < 0)) {
#pragma unroll
      for (int c = 0; c < 6; c++) {
        double v = w[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        w[c] = ((threadIdx.x & 63) == 0) ? v : 0.0;
      }
      sys = ((threadIdx.x & 63) == 0) ? s0 : -1;
    }
    add_out();
  }
};

}  // namespace mdg
