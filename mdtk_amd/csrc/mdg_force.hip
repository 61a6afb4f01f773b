// Force kernels of the REBO / AIREBO-LJ / torsion chain, templated on the
// geometry type R (double: the FP64 drop-in, reference operation order, CGS;
// float: the mixed-precision form, Angstrom/eV geometry with FP64 sums).
//
// Force scheme: the reference evaluates every energy sub-term once for its
// value and again for each derivative seed (REBO.cxx:200-814).  Here every
// bond (and every b* pair of the LJ term) runs one forward pass that keeps the
// scalar intermediates, then one adjoint pass seeded with dE/db that scatters
// V*dterm/dx into the gradient with FP64 atomics.
#include "mdg_kernels.hpp"

namespace mdg {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// LDS gradient-cache slots per workgroup of each scattering kernel.
constexpr int kTorsCache = 1024;
constexpr int kBondCache = 1024;
constexpr int kPairCache = 512;

// The workgroup's place in the work order.  The dispatcher deals
// workgroups to the eight XCDs in turn (workgroup b runs on XCD b mod 8);
// with map != 0 (and a grid that is a multiple of 8) XCD x takes the x-th
// contiguous eighth of the work, so that neighbouring work -- and the
// gradient lines it scatters into -- stays in one XCD's L2.
__device__ __forceinline__ int work_block(int map) {
  const int g = gridDim.x, b = blockIdx.x;
  if (map && (g & 7) == 0) return (b & 7) * (g >> 3) + (b >> 3);
  return b;
}

// The contiguous share [b0, b1) of nwork items of this workgroup, a
// multiple of the `unit` items it takes per pass (so that its gradient cache
// sees neighbouring work).
__device__ __forceinline__ void block_share(int nwork, int& b0, int& b1, int map, int unit = 256) {
  long long per = ((long long)(nwork + gridDim.x - 1) / gridDim.x + unit - 1) / unit * unit;
  long long lo = per * work_block(map);
  b0 = (int)min(lo, (long long)nwork);
  b1 = (int)min(lo + per, (long long)nwork);
}

// Parameter-gradient layout (per system, kNumTheta doubles): for the type
// pairs p = C-C, C-H, H-H (p = ti + tj), REBO's pair parameters at 9p + {Q, A,
// alpha, B1, B2, B3, beta1, beta2, beta3} (REBO.cxx:856-941) and the LJ
// parameters at 27 + 2p + {epsilon, sigma} (AIREBO_LJ.cxx:229-247).
__device__ __forceinline__ double* theta_of(double* dth, const Atoms& A, int owner) {
  return dth + (size_t)kNumTheta * A.sys[owner];
}

// Energy of one bond, pair or quadruplet, credited to its owner atom (the
// per-system totals are summed over atoms afterwards).
__device__ __forceinline__ void add_energy(const Atoms& A, double* eatom, int owner, double e) {
  energy_add(A, eatom, owner, e);
}

// Energy of a kernel unit expressed in erg.
template <class R>
__device__ __forceinline__ double to_erg(double e, double eunit) {
  if constexpr (std::is_same<R, double>::value) {
    return e;
  } else {
    return e * eunit;
  }
}

template <class R>
struct BondRef {
  int k;
  R r, f, df;
  V3<R> rv;  // x_a - x_k (minimum image)
  int tk;    // k's type and total coordination (LDS shells only)
  double Nk;
};

// The bond records of the kernel's type.
template <class R>
__device__ __forceinline__ BondRec<R>* brec(const Atoms& A) {
  if constexpr (std::is_same<R, double>::value) {
    return A.b64;
  } else {
    return A.b32;
  }
}

template <class R>
__device__ __forceinline__ BondRef<R> bond_at(const Atoms& A, int a, int s) {
  int o = a * A.maxb + s;
  BondRec<R> q = brec<R>(A)[o];
  BondRef<R> b;
  b.k = A.bj[o];
  b.r = q.r;
  b.f = q.f;
  b.df = q.df;
  b.rv = v3<R>(q.x, q.y, q.z);
  return b;
}

// The bond-order code reads each atom's shell -- its bond records, its
// coordination sums, and each partner's type and coordination -- through an
// accessor: GlobalShell reads global memory (the per-step bond lists), the
// LDS shells of k_rebo_shell read a wavefront's staged copy.  view(a)
// resolves an atom once; nb/bond/Nt/NH/NC read through the view.
struct GView {
  int a;
};
template <class R>
struct GlobalShell {
  const Atoms& A;
  __device__ __forceinline__ GView view(int a) const { return GView{a}; }
  __device__ __forceinline__ int nb(GView v) const { return A.bcnt[v.a]; }
  __device__ __forceinline__ BondRef<R> bond(GView v, int t) const { return bond_at<R>(A, v.a, t); }
  __device__ __forceinline__ double Nt(GView v) const { return A.Nt[v.a]; }
  __device__ __forceinline__ double NH(GView v) const { return A.NH[v.a]; }
  __device__ __forceinline__ double NC(GView v) const { return A.NC[v.a]; }
  // a bond partner's type and total coordination
  __device__ __forceinline__ int ktype(const BondRef<R>& k) const { return A.type[k.k]; }
  __device__ __forceinline__ double kNt(const BondRef<R>& k) const { return A.Nt[k.k]; }
};

// A wavefront's staged shells (k_rebo_shell): the atoms of its tile of
// bonds, each with its coordination sums and its bond records, every record
// carrying the partner's index, type and total coordination, so that the
// bond order reads LDS where the global form gathers.  An atom that did not
// fit (more than kShAtoms atoms or kShRecs records in the tile) is read
// from global memory.
constexpr int kShAtoms = 96;   // atoms per wave
constexpr int kShRecs = 352;   // bond records per wave
constexpr int kShHash = 256;   // hash slots per wave (power of two)
template <class R>
struct ShellRec {
  R x, y, z, r, f, df, Nk;
  unsigned kt;  // partner index << 2 | partner type
};
struct LView {
  int a, s;  // atom, staged slot (-1: global)
};
__device__ __forceinline__ int shell_hash(int a) { return (int)(((unsigned)a * 2654435761u) >> 24); }
template <class R>
struct LdsShell {
  const Atoms& A;
  const MDG_LDS int* key;    // [kShHash] atom, -1 free
  const MDG_LDS short* slot; // [kShHash] slot of the atom, -1 not staged
  const MDG_LDS int* nbs;    // [kShAtoms] bonds, -1 when the records did not fit
  const MDG_LDS int* off;    // [kShAtoms] first record
  const MDG_LDS double* N;   // [kShAtoms][3] Nt, NH, NC
  const MDG_LDS ShellRec<R>* rec;
  __device__ __forceinline__ LView view(int a) const {
    int h = shell_hash(a);
#pragma unroll 1
    for (int p = 0; p < kShHash; p++) {
      const int k = key[h];
      if (k == a) {
        const int sl = slot[h];
        return LView{a, (sl >= 0 && nbs[sl] >= 0) ? sl : -1};
      }
      if (k == -1) break;
      h = (h + 1) & (kShHash - 1);
    }
    return LView{a, -1};
  }
  __device__ __forceinline__ int nb(LView v) const { return v.s >= 0 ? nbs[v.s] : A.bcnt[v.a]; }
  __device__ __forceinline__ BondRef<R> bond(LView v, int t) const {
    BondRef<R> b;
    if (v.s >= 0) {
      const MDG_LDS ShellRec<R>* q = rec + off[v.s] + t;
      const unsigned kt = q->kt;
      b.k = (int)(kt >> 2);
      b.tk = (int)(kt & 3u);
      b.r = q->r;
      b.f = q->f;
      b.df = q->df;
      b.rv = v3<R>(q->x, q->y, q->z);
      b.Nk = double(q->Nk);
    } else {
      b = bond_at<R>(A, v.a, t);
      b.tk = A.type[b.k];
      b.Nk = A.Nt[b.k];
    }
    return b;
  }
  __device__ __forceinline__ double Nt(LView v) const { return v.s >= 0 ? N[3 * v.s] : A.Nt[v.a]; }
  __device__ __forceinline__ double NH(LView v) const { return v.s >= 0 ? N[3 * v.s + 1] : A.NH[v.a]; }
  __device__ __forceinline__ double NC(LView v) const { return v.s >= 0 ? N[3 * v.s + 2] : A.NC[v.a]; }
  __device__ __forceinline__ int ktype(const BondRef<R>& k) const { return k.tk; }
  __device__ __forceinline__ double kNt(const BondRef<R>& k) const { return k.Nk; }
};

// Coordination of a excluding b (REBO::Nt/NH/NC, REBO.cxx:423-493): the
// cached FP64 sum minus f_ab when r_ab < R2(a,b), clamped at 0.
template <class R>
__device__ __forceinline__ double coord_excl(double Nsum, bool subtract, R r, R R2, R fab, unsigned* err) {
  double N = Nsum;
  if (subtract && r < R2) N -= double(fab);
  if (!(N > -1e-15)) raise(err, ERR_NEG_COORD);
  if (N < 0) N = 0;
  return N;
}
}  // namespace

// ----------------------------------------------------- coordination pass
// REBO::countNeighbours (REBO.cxx:328-393) plus the per-atom bond list: all
// short-list neighbours with r <= R2 in list order, with r, f, df and the
// minimum-image vector.  A group of 16 lanes per atom takes 16 row entries
// per round (coalesced row reads and record writes); bonds keep their row
// order by ballot compaction, and Nt/NH/NC are summed in that order (the
// reference's), one shuffle per bond.
template <class R>
__global__ void __launch_bounds__(256) k_coord(Atoms A, const TablesT<R>* __restrict__ T, Box bx, unsigned* err) {
  constexpr int LW = 16;
  const int lane = threadIdx.x & 63;
  const int sl = lane % LW;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mine = ((1ull << LW) - 1ull) << (lane / LW * LW);
  const int i = blockIdx.x * (blockDim.x / LW) + threadIdx.x / LW;
  if (i >= A.n) return;  // group-uniform
  BondRec<R>* B = brec<R>(A);
  const int ti = A.type[i];
  unsigned* const er = sys_err(err, A.sys[i]);
  double Nt = 0, NH = 0, NC = 0;
  float eH[2] = {0.0f, 0.0f};  // an H centre's sum of f ExpTerm for a partner of type C, H (b* bound)
  int nb = 0;
  if (ti <= T_H) {
    const int cnt = A.nls_cnt[i];
    const int* L = A.nls + (size_t)i * A.cap_s;
    const double4 pi = A.p[i];
    const Frame fr = frame_of(bx, i);
    // the centre's cutoffs and H ExpTerm densities for partners C and H,
    // loaded once per atom and selected by the partner's type
    const R rlo[2] = {T->rebo.R[ti][T_C][0], T->rebo.R[ti][T_H][0]};
    const R rhi[2] = {T->rebo.R[ti][T_C][1], T->rebo.R[ti][T_H][1]};
    const float rhoH[2] = {float(T->rebo.rho[T_H][T_C]), float(T->rebo.rho[T_H][T_H])};
    for (int b0 = 0; b0 < cnt; b0 += LW) {
      const int q = b0 + sl;
      bool ok = false;
      int k = 0, tk = 0;
      V3<R> rv = v3<R>(0, 0, 0);
      R r = R(0), f = R(0), df = R(0);
      if (q < cnt) {
        k = L[q];
        const double4 pk = A.p[k];
        tk = (int)pk.w;  // the packed position carries the type
        if ((unsigned)tk > (unsigned)T_OTHER) raise(er, ERR_TYPE_RANGE);
        rv = to_local<R>(disp(bx, fr, i, k, pi, pk));
        r = sqrt(rv.x * rv.x + rv.y * rv.y + rv.z * rv.z);
        if ((unsigned)tk <= (unsigned)T_H) {
          const R lo = (tk == T_C) ? rlo[0] : rlo[1], hi = (tk == T_C) ? rhi[0] : rhi[1];
          if (!(r > hi)) {
            ok = true;
            cutoff<R>(r, lo, hi, f, df);
          }
        }
      }
      // an H centre's f ExpTerm for a b* partner of type C, H (r_alter = R0,
      // REBO.cxx:221-226), one bond per lane
      float ex[2] = {0.0f, 0.0f};
      if (ok && ti == T_H && A.pmax) {
        const ReboParamsT<R>& P = T->rebo;
#pragma unroll
        for (int tb = 0; tb < 2; tb++)
          ex[tb] = float(f) * expf(float(P.expscale) * ((((tk == T_C) ? rhoH[0] : rhoH[1]) - float(r)) -
                                                        (float(P.rho[T_H][tb]) - float(P.R[T_H][tb][0]))));
      }
      const unsigned long long m = __ballot(ok) & mine;
      if (ok) {
        const int slot = nb + __popcll(m & below);
        if (slot < A.maxb) {
          const int o = i * A.maxb + slot;
          A.bj[o] = k;
          B[o] = BondRec<R>{rv.x, rv.y, rv.z, r, f, df, R(0), R(0)};
        }
      }
      for (unsigned long long mm = m; mm; mm &= mm - 1) {  // in row order
        const int src = __ffsll((long long)mm) - 1;
        const double fs = __shfl(double(f), src, 64);
        const int ts = __shfl(tk, src, 64);
        Nt += fs;
        if (ts == T_H) NH += fs;
        if (ts == T_C) NC += fs;
        if (ti == T_H && A.pmax) {  // group-uniform
          eH[0] += __shfl(ex[0], src, 64);
          eH[1] += __shfl(ex[1], src, 64);
        }
      }
      nb += __popcll(m);
    }
  }
  if (sl == 0) {
    if (nb > A.maxb) raise(er, ERR_BOND_OVERFLOW);
    if (A.pmax) {
      // D_a = 1 + sum_k f G E (+ P for a C centre) over a's bonds, partner
      // not among them: G >= glb, E = 1 for C; P >= plb.  In FP32 with
      // relative margins of 1e-5 (the table bounds' own margins are 1e-4)
#pragma unroll
      for (int tb = 0; tb < 2; tb++) {
        float D = 0.0f;
        if (ti == T_C) D = 1.0f + float(T->plb[tb]) + float(T->glb[0]) * float(Nt);
        else if (ti == T_H) D = 1.0f + float(T->glb[1]) * eH[tb];
        D = D * (1.0f - 1e-5f) - 1e-5f;
        A.pmax[2 * i + tb] = (D > 0.0f) ? rsqrtf(D) * (1.0f + 1e-5f) : 1e30f;
      }
    }
    A.Nt[i] = Nt;
    A.NH[i] = NH;
    A.NC[i] = NC;
    A.bcnt[i] = nb < A.maxb ? nb : A.maxb;
  }
}

// ------------------------------------------------------- bond order b_ij
// The pair (i,j) whose bond order is needed: a REBO bond, or an LJ pair for
// b* (then r_eff = R0 inside ExpTerm only, AIREBO_LJ.cxx:191-202).
template <class R>
struct BPair {
  int i, j, ti, tj;
  V3<R> rv;  // x_i - x_j
  R r;
  R f;       // REBO cutoff value of (i,j) (0 for b* pairs beyond R2)
  R reff;
};

struct BOState {
  double D[2];        // D_ij, D_ji
  double N[2];        // Nt_ij, Nt_ji
  double NHc[2], NCc[2];
  double P[2][3];     // P and dP/dNH, dP/dNC per side
  double sum[2];      // Nconj partial sums (i side, j side)
  double Nconj;
  double rc[4];       // pi_rc and d/dNi, d/dNj, d/dNconj in (i,j) orientation
  double T[4];        // T_ij and derivatives
  double dh;          // pi_dh angular sum
  double b;
};

// Centre a, partner b of side s (0: centre i; 1: centre j).
template <class R>
struct Side {
  int a, b, ta, tb;
  V3<R> rv;  // x_a - x_b
};
template <class R>
__device__ __forceinline__ Side<R> side_of(const BPair<R>& p, int s) {
  Side<R> q;
  if (s == 0) {
    q.a = p.i; q.b = p.j; q.ta = p.ti; q.tb = p.tj; q.rv = p.rv;
  } else {
    q.a = p.j; q.b = p.i; q.ta = p.tj; q.tb = p.ti; q.rv = -p.rv;
  }
  return q;
}

// The REBO scalar parameters and G splines of a bond-order kernel's
// workgroup, copied into LDS once (stage_rebo, before the kernel's first
// barrier) and read from there by the bond-order code: their indices vary
// per lane, so from global memory every read is a gather.
__shared__ ReboParamsT<float> s_rebo_f;
__shared__ ReboParamsT<double> s_rebo_d;
template <class R>
__device__ __forceinline__ const ReboParamsT<R>& rebo_lds() {
  if constexpr (std::is_same<R, float>::value) return s_rebo_f;
  else return s_rebo_d;
}
template <class R>
__device__ __forceinline__ void stage_rebo(const TablesT<R>* __restrict__ T) {
  static_assert(sizeof(ReboParamsT<R>) % 4 == 0, "copied as words");
  const unsigned* src = reinterpret_cast<const unsigned*>(&T->rebo);
  unsigned* dst = reinterpret_cast<unsigned*>(const_cast<ReboParamsT<R>*>(&rebo_lds<R>()));
  for (int w = threadIdx.x; w < (int)(sizeof(ReboParamsT<R>) / 4); w += blockDim.x) dst[w] = src[w];
}

// G(cos theta_bak) with its partial derivatives (REBO.cxx:277-326).
template <class R>
struct GVal {
  R G, dGdc, dGdN;
  bool has_dc, has_dN;
};
template <class R>
__device__ __forceinline__ GVal<R> gfun(const ReboParamsT<R>& P, int ta, R c, R N) {
  GVal<R> g;
  if (ta == T_C) {
    R Sp = sprime<R>(N, P.Ntot[0], P.Ntot[1]);
    R g1, g1d, g2, g2d;
    gspline<R>(P.gC1, c, g1, g1d);
    gspline<R>(P.gC2, c, g2, g2d);
    g.G = g2 + Sp * (g1 - g2);
    g.has_dc = (g2d != R(0) || g1d != R(0));
    g.dGdc = g2d + g1d * Sp + (-g2d * Sp);
    R dSp = dsprime<R>(N, P.Ntot[0], P.Ntot[1]);
    g.has_dN = ((g2 != R(0) || g1 != R(0)) && dSp != R(0));
    g.dGdN = g1 * dSp + (-g2 * dSp);
  } else {
    R v, d;
    gspline<R>(P.gH, c, v, d);
    g.G = v;
    g.has_dc = (d != R(0));
    g.dGdc = d;
    g.has_dN = false;
    g.dGdN = R(0);
  }
  return g;
}

template <class R>
__device__ __forceinline__ R exp_term(const ReboParamsT<R>& P, const Side<R>& q, int tk, R r_ak, R reff_ab) {
  if (q.ta != T_H) return R(1);
  return exp(P.expscale * ((P.rho[q.ta][tk] - r_ak) - (P.rho[q.ta][q.tb] - reff_ab)));
}

template <class R>
__device__ __forceinline__ R cos_theta(V3<R> a, V3<R> b, R ra, R rb, unsigned* err, R& raw) {
  R e1 = dot(a, b);
  R e2 = ra * rb;
  R c = e1 / e2;
  raw = c;
  // FGeneral.hpp:147 allows 1e-7 beyond +-1 in FP64; FP32 rounding of a
  // near-collinear triplet alone reaches a few ulps (1.2e-7 each), so the
  // FP32 form allows 32 ulps.
  constexpr double tol = std::is_same<R, double>::value ? 1e-7 : 32.0 * 1.1920929e-7;
  if (!(c >= R(-1.0 - tol) && c <= R(1.0 + tol))) raise(err, ERR_COS_RANGE);
  if (c < R(-1)) c = R(-1);
  if (c > R(1)) c = R(1);
  return c;
}

// First part of the forward pass of b_ij: the coordinations excluding the
// pair and D_ij, D_ji of the sigma-pi terms (REBO.cxx:235-262, P :626-657).
template <class R, class SH>
__device__ __forceinline__ void bo_sigma_pi(const SH& sh, const TablesT<R>* __restrict__ T, const BPair<R>& p,
                                            BOState& st, unsigned* err) {
  const ReboParamsT<R>& P = rebo_lds<R>();
  R R2ij = P.R[p.ti][p.tj][1];
  const auto vi = sh.view(p.i), vj = sh.view(p.j);
  st.N[0] = coord_excl<R>(sh.Nt(vi), true, p.r, R2ij, p.f, err);
  st.N[1] = coord_excl<R>(sh.Nt(vj), true, p.r, R2ij, p.f, err);
  st.NHc[0] = coord_excl<R>(sh.NH(vi), p.tj == T_H, p.r, R2ij, p.f, err);
  st.NCc[0] = coord_excl<R>(sh.NC(vi), p.tj == T_C, p.r, R2ij, p.f, err);
  st.NHc[1] = coord_excl<R>(sh.NH(vj), p.ti == T_H, p.r, R2ij, p.f, err);
  st.NCc[1] = coord_excl<R>(sh.NC(vj), p.ti == T_C, p.r, R2ij, p.f, err);
#pragma unroll
  for (int s = 0; s < 2; s++) {
    Side<R> q = side_of(p, s);
    double D = 1.0;
    const auto va = sh.view(q.a);
    int nb = sh.nb(va);
    for (int t = 0; t < nb; t++) {
      BondRef<R> k = sh.bond(va, t);
      if (k.k == q.b) continue;
      int tk = sh.ktype(k);
      R E = exp_term<R>(P, q, tk, k.r, p.reff);
      if (E != R(0)) {
        R raw;
        R c = cos_theta<R>(q.rv, k.rv, p.r, k.r, err, raw);
        GVal<R> g = gfun<R>(P, q.ta, c, R(st.N[s]));
        D += double(g.G * k.f * E);
      }
    }
    st.P[s][0] = st.P[s][1] = st.P[s][2] = 0.0;
    if (q.ta == T_C && q.tb <= T_H) {
      R o[3];
      bicubic<R>(T->P[q.tb == T_C ? 0 : 1], R(st.NHc[s]), R(st.NCc[s]), o);
      st.P[s][0] = o[0];
      st.P[s][1] = o[1];
      st.P[s][2] = o[2];
    }
    D += st.P[s][0];
    st.D[s] = D;
  }
}

// The rest of the forward pass of b_ij = (p_ij + p_ji)/2 + pi_rc + pi_dh
// (REBO::Baver, REBO.cxx:200-211, with p_sigma_pi :264-275, Nconj :540-624,
// pi_rc :659-714, pi_dh :716-814), after bo_sigma_pi.
template <class R, class SH>
__device__ __forceinline__ void bo_finish(const SH& sh, const TablesT<R>* __restrict__ T, const BPair<R>& p,
                                          BOState& st, unsigned* err) {
  const ReboParamsT<R>& P = rebo_lds<R>();
  // D^-1/2: the library pow as the reference (FP64 form), 1/sqrt (mixed)
  double pij, pji;
  if constexpr (std::is_same<R, float>::value) {
    pij = 1.0 / sqrt(st.D[0]);
    pji = 1.0 / sqrt(st.D[1]);
  } else {
    pij = pow(st.D[0], -0.5);
    pji = pow(st.D[1], -0.5);
  }
  // Nconj partial sums: carbon neighbours k of the centre, f_ka * S'(N_ka)
#pragma unroll
  for (int s = 0; s < 2; s++) {
    Side<R> q = side_of(p, s);
    double sum = 0.0;
    const auto va = sh.view(q.a);
    int nb = sh.nb(va);
    for (int t = 0; t < nb; t++) {
      BondRef<R> k = sh.bond(va, t);
      if (k.k == q.b) continue;
      int tk = sh.ktype(k);
      if (tk != T_C) continue;
      double Nk = coord_excl<R>(sh.kNt(k), true, k.r, P.R[tk][q.ta][1], k.f, err);
      sum += double(k.f * sprime<R>(R(Nk), P.Ntot[0], P.Ntot[1]));
    }
    st.sum[s] = sum;
  }
  st.Nconj = 1.0 + st.sum[0] * st.sum[0] + st.sum[1] * st.sum[1];
  // pi_rc by pair type; the H-C case evaluates F_CH(N_j, N_i, Nconj_ji)
  for (int q = 0; q < 4; q++) st.rc[q] = 0.0;
  R o[4];
  if (p.ti == T_C && p.tj == T_C) {
    tricubic<R, 16>(T->T3[0], R(st.N[0]), R(st.N[1]), R(st.Nconj), 15, 15, o);
    for (int q = 0; q < 4; q++) st.rc[q] = o[q];
  } else if (p.ti == T_C && p.tj == T_H) {
    tricubic<R, 16>(T->T3[1], R(st.N[0]), R(st.N[1]), R(st.Nconj), 15, 15, o);
    for (int q = 0; q < 4; q++) st.rc[q] = o[q];
  } else if (p.ti == T_H && p.tj == T_C) {
    double nconj_ji = 1.0 + st.sum[1] * st.sum[1] + st.sum[0] * st.sum[0];
    tricubic<R, 16>(T->T3[1], R(st.N[1]), R(st.N[0]), R(nconj_ji), 15, 15, o);
    st.rc[0] = o[0];
    st.rc[1] = o[2];  // d/dN_i
    st.rc[2] = o[1];  // d/dN_j
    st.rc[3] = o[3];
  } else if (p.ti == T_H && p.tj == T_H) {
    tricubic<R, 16>(T->T3[2], R(st.N[0]), R(st.N[1]), R(st.Nconj), 15, 15, o);
    for (int q = 0; q < 4; q++) st.rc[q] = o[q];
  }
  // pi_dh: C-C only, skipped when T_ij == 0 (REBO.cxx:802-814)
  st.T[0] = st.T[1] = st.T[2] = st.T[3] = 0.0;
  st.dh = 0.0;
  double pidh = 0.0;
  if (p.ti == T_C && p.tj == T_C) {
    tricubic<R, 16>(T->T3[3], R(st.N[0]), R(st.N[1]), R(st.Nconj), 15, 15, o);
    for (int q = 0; q < 4; q++) st.T[q] = o[q];
    if (st.T[0] != 0.0) {
      double s = 0.0;
      const auto vi = sh.view(p.i), vj = sh.view(p.j);
      int nbi = sh.nb(vi), nbj = sh.nb(vj);
      for (int a = 0; a < nbi; a++) {
        BondRef<R> k = sh.bond(vi, a);
        if (k.k == p.j) continue;
        for (int b = 0; b < nbj; b++) {
          BondRef<R> l = sh.bond(vj, b);
          if (l.k == p.i) continue;
          if (k.k == l.k) continue;
          R sin_ik = sqrt(norm2(cross(p.rv, k.rv))) / (p.r * k.r);
          if (fabs(sin_ik) < R(0.1)) continue;
          int tk = sh.ktype(k), tl = sh.ktype(l);
          R fik = sprime<R>(k.r, P.Rp[tk][p.ti][0], P.Rp[tk][p.ti][1]);
          R sin_jl = sqrt(norm2(cross(p.rv, l.rv))) / (p.r * l.r);
          if (fabs(sin_jl) < R(0.1)) continue;
          R fjl = sprime<R>(l.r, P.Rp[tl][p.tj][0], P.Rp[tl][p.tj][1]);
          V3<R> Pp = cross(-p.rv, k.rv), Q = cross(p.rv, l.rv);
          R e2 = sqrt(norm2(Pp)) * sqrt(norm2(Q));
          R cd = (e2 == R(0)) ? R(0) : dot(Pp, Q) / e2;
          if (cd < R(-1)) cd = R(-1);
          if (cd > R(1)) cd = R(1);
          R c = -cd;
          s += double((R(1) - c * c) * fik * fjl);
        }
      }
      st.dh = s;
      pidh = st.T[0] * s;
    }
  }
  st.b = (pij + pji) / 2.0 + st.rc[0] + pidh;
}

template <class R, class SH>
__device__ __forceinline__ void bo_forward(const SH& sh, const TablesT<R>* __restrict__ T, const BPair<R>& p,
                                           BOState& st, unsigned* err) {
  bo_sigma_pi<R>(sh, T, p, st, err);
  bo_finish<R>(sh, T, p, st, err);
}

// Scatter c * d f_ak / dx for the pair (a,k) along u_ak = rv/r (a +=, k -=).
template <class R, class G, class W>
__device__ __forceinline__ void pair_push(const G& g, int k, V3<R> rv, R r, R c, V3<R>& ga, double gs, W& vw) {
  V3<R> v = rv * (c / r);
  ga = ga + v;
  vw.push(rv, v);
  scatter3<R>(g, k, -v, gs);
}

// Adjoint pass: adds W * db_ij/dx to grad (W in the kernel's energy unit);
// every contribution is a push +v, -v along one of the term's pair vectors,
// which vw adds to the virial.
template <class R, class SH, class G, class Wv>
__device__ __forceinline__ void bo_backward(const SH& sh, const TablesT<R>* __restrict__ T, const BPair<R>& p,
                                            const BOState& st, double W, G grad, double gs, unsigned* err,
                                            Wv& vw) {
  const ReboParamsT<R>& P = rebo_lds<R>();
  V3<R> gi = v3<R>(0, 0, 0), gj = v3<R>(0, 0, 0);
  double dN[2] = {0, 0}, dNH[2] = {0, 0}, dNC[2] = {0, 0}, dNconj = 0.0;
  // sigma-pi terms
#pragma unroll
  for (int s = 0; s < 2; s++) {
    Side<R> q = side_of(p, s);
    V3<R>& ga = (s == 0) ? gi : gj;
    V3<R>& gb = (s == 0) ? gj : gi;
    double seedD;
    if constexpr (std::is_same<R, float>::value) {
      const double q = 1.0 / sqrt(st.D[s]);
      seedD = (-0.5 * (q * q * q)) * (W / 2.0);
    } else {
      seedD = (-0.5 * pow(st.D[s], -0.5 - 1.0)) * (W / 2.0);
    }
    R seedR = R(seedD);
    const auto va = sh.view(q.a);
    int nb = sh.nb(va);
    for (int t = 0; t < nb; t++) {
      BondRef<R> k = sh.bond(va, t);
      if (k.k == q.b) continue;
      int tk = sh.ktype(k);
      R E = exp_term<R>(P, q, tk, k.r, p.reff);
      if (E == R(0)) continue;
      R raw;
      R c = cos_theta<R>(q.rv, k.rv, p.r, k.r, err, raw);
      GVal<R> g = gfun<R>(P, q.ta, c, R(st.N[s]));
      R V1 = k.f * E * seedR;
      V3<R> gk = v3<R>(0, 0, 0);
      if (g.has_dc) {
        R cc = g.dGdc * V1;
        R e2 = p.r * k.r;
        V3<R> da = k.rv / e2 - q.rv * (raw / (p.r * p.r));
        V3<R> db = q.rv / e2 - k.rv * (raw / (k.r * k.r));
        ga = ga + (da + db) * cc;
        gb = gb - da * cc;
        gk = gk - db * cc;
        vw.push(q.rv, da * cc);
        vw.push(k.rv, db * cc);
      }
      if (g.has_dN) dN[s] += double(g.dGdN * V1);
      if (k.df != R(0)) {
        V3<R> v = k.rv * (k.df * (g.G * E * seedR) / k.r);
        ga = ga + v;
        gk = gk - v;
        vw.push(k.rv, v);
      }
      if (q.ta == T_H) {
        R dt = E * P.expscale;
        R V3s = g.G * k.f * seedR;
        // r_ab (or the b* pair's r_alter, whose derivative is the unit vector)
        V3<R> u = q.rv * (dt * V3s / p.r);
        ga = ga + u;
        gb = gb - u;
        V3<R> v = k.rv * (-dt * V3s / k.r);
        ga = ga + v;
        gk = gk - v;
        vw.push(q.rv, u);
        vw.push(k.rv, v);
      }
      scatter3<R>(grad, k.k, gk, gs);
    }
    dNH[s] += st.P[s][1] * seedD;
    dNC[s] += st.P[s][2] * seedD;
  }
  // pi_rc
  dN[0] += st.rc[1] * W;
  dN[1] += st.rc[2] * W;
  dNconj += st.rc[3] * W;
  // pi_dh
  if (st.T[0] != 0.0) {
    double Wt = st.dh * W;
    if (st.dh != 0.0) {
      dN[0] += st.T[1] * Wt;
      dN[1] += st.T[2] * Wt;
      dNconj += st.T[3] * Wt;
    }
    R Ws = R(st.T[0] * W);
    const auto vi = sh.view(p.i), vj = sh.view(p.j);
      int nbi = sh.nb(vi), nbj = sh.nb(vj);
    for (int a = 0; a < nbi; a++) {
      BondRef<R> k = sh.bond(vi, a);
      if (k.k == p.j) continue;
      for (int b = 0; b < nbj; b++) {
        BondRef<R> l = sh.bond(vj, b);
        if (l.k == p.i) continue;
        if (k.k == l.k) continue;
        R sin_ik = sqrt(norm2(cross(p.rv, k.rv))) / (p.r * k.r);
        if (fabs(sin_ik) < R(0.1)) continue;
        int tk = sh.ktype(k), tl = sh.ktype(l);
        R R1k = P.Rp[tk][p.ti][0], R2k = P.Rp[tk][p.ti][1];
        R fik = sprime<R>(k.r, R1k, R2k);
        R sin_jl = sqrt(norm2(cross(p.rv, l.rv))) / (p.r * l.r);
        if (fabs(sin_jl) < R(0.1)) continue;
        R R1l = P.Rp[tl][p.tj][0], R2l = P.Rp[tl][p.tj][1];
        R fjl = sprime<R>(l.r, R1l, R2l);
        V3<R> a_ = p.rv, b_ = k.rv, c_ = l.rv;
        V3<R> Pp = cross(-a_, b_), Q = cross(a_, c_);
        R p2 = norm2(Pp), q2 = norm2(Q);
        R e2 = sqrt(p2) * sqrt(q2);
        R cs = (e2 == R(0)) ? R(0) : dot(Pp, Q) / e2;
        R cd = cs;
        if (cd < R(-1)) cd = R(-1);
        if (cd > R(1)) cd = R(1);
        R c = -cd;
        V3<R> gk = v3<R>(0, 0, 0), gl = v3<R>(0, 0, 0);
        R V = R(2) * c * fik * fjl * Ws;
        if (V != R(0) && e2 != R(0)) {
          V3<R> gP = Q / e2 - Pp * (cs / p2);
          V3<R> gQ = Pp / e2 - Q * (cs / q2);
          V3<R> da = cross(gP, b_) + cross(c_, gQ);
          V3<R> db = cross(a_, gP);
          V3<R> dc = cross(gQ, a_);
          gi = gi + (da + db) * V;
          gj = gj + (dc - da) * V;
          gk = gk - db * V;
          gl = gl - dc * V;
          vw.push(a_, da * V);
          vw.push(b_, db * V);
          vw.push(c_, dc * V);
        }
        R Vk = (R(1) - c * c) * fjl * Ws;
        if (Vk != R(0) && k.r >= R1k && k.r <= R2k) {
          V3<R> v = k.rv * (dsprime<R>(k.r, R1k, R2k) * Vk / k.r);
          gi = gi + v;
          gk = gk - v;
          vw.push(k.rv, v);
        }
        R Vl = (R(1) - c * c) * fik * Ws;
        if (Vl != R(0) && l.r >= R1l && l.r <= R2l) {
          V3<R> v = l.rv * (dsprime<R>(l.r, R1l, R2l) * Vl / l.r);
          gj = gj + v;
          gl = gl - v;
          vw.push(l.rv, v);
        }
        scatter3<R>(grad, k.k, gk, gs);
        scatter3<R>(grad, l.k, gl, gs);
      }
    }
  }
  // coordination derivatives (dfThem, REBO.cxx:408-421) and Nconj
#pragma unroll
  for (int s = 0; s < 2; s++) {
    Side<R> q = side_of(p, s);
    V3<R>& ga = (s == 0) ? gi : gj;
    double sconj = (st.sum[s] != 0.0) ? 2.0 * st.sum[s] * dNconj : 0.0;
    const auto va = sh.view(q.a);
    int nb = sh.nb(va);
    for (int t = 0; t < nb; t++) {
      BondRef<R> k = sh.bond(va, t);
      if (k.k == q.b) continue;
      int tk = sh.ktype(k);
      double c = dN[s];
      if (tk == T_H) c += dNH[s];
      if (tk == T_C) c += dNC[s];
      double coefk = 0.0;
      if (tk == T_C && sconj != 0.0) {
        double Nk = coord_excl<R>(sh.kNt(k), true, k.r, P.R[tk][q.ta][1], k.f, err);
        double Sp = sprime<R>(R(Nk), P.Ntot[0], P.Ntot[1]);
        c += Sp * sconj;
        coefk = double(k.f) * double(dsprime<R>(R(Nk), P.Ntot[0], P.Ntot[1])) * sconj;
      }
      if (c != 0.0 && k.df != R(0)) pair_push<R>(grad, k.k, k.rv, k.r, R(double(k.df) * c), ga, gs, vw);
      if (coefk != 0.0) {
        // d N_ka: neighbours m of k other than a
        V3<R> gk = v3<R>(0, 0, 0);
        const auto vk = sh.view(k.k);
        int nbk = sh.nb(vk);
        for (int u = 0; u < nbk; u++) {
          BondRef<R> m = sh.bond(vk, u);
          if (m.k == q.a || m.df == R(0)) continue;
          V3<R> v = m.rv * R(double(m.df) * coefk / double(m.r));
          gk = gk + v;
          vw.push(m.rv, v);
          scatter3<R>(grad, m.k, -v, gs);
        }
        scatter3<R>(grad, k.k, gk, gs);
      }
    }
  }
  scatter3<R>(grad, p.i, gi, gs);
  scatter3<R>(grad, p.j, gj, gs);
}

// ------------------------------------------------------------ REBO bonds
// REBO::operator() (REBO.cxx:32-65): E += VR + b_ij * VA for every bond.
// The pair terms VR, VA and their radial derivatives are evaluated in FP64
// from the FP64 table T64 in both forms: near equilibrium they are two
// ~100 eV/A forces that cancel, so FP32's 1e-7 relative rounding would leave
// ~1e-5 eV/A per bond.  The bond order b_ij (and its adjoint) runs in R.
// TH: also accumulate dE/dtheta of the pair parameters into dth (FP64 only).
// OCC: waves per SIMD the register allocation is held to (a tuning choice).
// VIR: also accumulate the per-system virial into vir (erg).
template <class R, bool TH, int OCC, bool VIR>
__global__ void __launch_bounds__(256, OCC) k_rebo_bonds(Atoms A, const TablesT<R>* __restrict__ T,
                                                   const Tables* __restrict__ T64, Box bx,
                                                   const int2* __restrict__ bonds, const int* __restrict__ nbonds,
                                                   double* __restrict__ grad, double* __restrict__ eatom,
                                                   unsigned* err, double* __restrict__ dth,
                                                   long long* __restrict__ vir) {
  using GC = CacheOf<R, kBondCache>;
  __shared__ int s_key[GC::kSlots];
  __shared__ typename GC::val_t s_val[GC::kVals];
  __shared__ int s_fill;
  GC gc{s_key, s_val, &s_fill, grad, A.gfix};
  gc.clear();
  stage_rebo<R>(T);
  __syncthreads();
  const double gs = T->eunit / T->lunit;
  const double su = 1.0 / T->eunit;  // erg -> the table's energy unit (adjoint seed)
  Virial<VIR> vw(vir, T->eunit);
  int nbt = min(*nbonds, A.bond_cap);
  int b0, b1;
  block_share(nbt, b0, b1, A.wg_map);
  for (int base = b0; base < b1; base += 256) {
    const int t = base + threadIdx.x;
    const ReboParamsT<double>& P = T64->rebo;
    int2 bd = (t < b1) ? bonds[t] : make_int2(0, 0);
    BPair<R> p;
    p.i = bd.x;
    p.j = bd.y;
    p.ti = A.type[p.i];
    p.tj = A.type[p.j];
    D3 rd = depos(A.p, p.i, p.j, bx);
    p.rv = to_local<R>(rd);
    p.r = sqrt(norm2(p.rv));
    int a = p.ti, b = p.tj;
    double r = sqrt(norm2(rd));
    double f, df;
    cutoff<double, std::is_same<R, float>::value>(r, P.R[a][b][0], P.R[a][b][1], f, df);
    if (t >= b1) f = 0.0;  // past the share: no work this pass
    if constexpr (VIR) {
      if (t < b1) vw.at(A.sys[p.i]);
    }
    unsigned* const er = sys_err(err, A.sys[p.i]);
    {  // f_ij inside the bond order in R, as the coordination pass summed it
      R fr, dfr;
      cutoff<R>(p.r, T->rebo.R[a][b][0], T->rebo.R[a][b][1], fr, dfr);
      p.f = fr;
    }
    p.reff = p.r;
    // VR (REBO.cxx:132-160) and VA (:162-198)
    if (f != 0.0) {
      double cr = 0.0;
      double t1 = P.A[a][b] * exp(-P.alpha[a][b] * r);
      double val = (1.0 + P.Q[a][b] / r) * t1;
      double der = -t1 * (P.Q[a][b] + P.alpha[a][b] * r * r + P.alpha[a][b] * r * P.Q[a][b]) / (r * r);
      double VR = f * val;
      cr += der * f + val * df;
      double u1 = P.B[0][a][b] * exp(-P.beta[0][a][b] * r);
      double u2 = P.B[1][a][b] * exp(-P.beta[1][a][b] * r);
      double u3 = P.B[2][a][b] * exp(-P.beta[2][a][b] * r);
      double vsum = u1 + u2 + u3;
      double VA = -f * vsum;
      double eb = VR;
      double bb = 0.0;
      if (VA != 0.0 || TH) {
        BOState st;
        bo_forward<R>(GlobalShell<R>{A}, T, p, st, er);
        bb = st.b;
        if (VA != 0.0) {
          bo_backward<R>(GlobalShell<R>{A}, T, p, st, VA * su, gc, gs, er, vw);
          double vder = (-P.beta[0][a][b]) * u1 + (-P.beta[1][a][b]) * u2 + (-P.beta[2][a][b]) * u3;
          cr += (-vder * f * bb) + (-vsum * bb) * df;
          eb += bb * VA;
        }
      }
      if constexpr (TH) {
        // VR = f (1 + Q/r) A e^{-alpha r}, b VA = -b f sum_n B_n e^{-beta_n r}
        double* th = theta_of(dth, A, p.i) + 9 * (a + b);
        double ea = exp(-P.alpha[a][b] * r);
        unsafeAtomicAdd(th + 0, f * P.A[a][b] * ea / r);
        unsafeAtomicAdd(th + 1, f * (1.0 + P.Q[a][b] / r) * ea);
        unsafeAtomicAdd(th + 2, -r * VR);
#pragma unroll
        for (int n = 0; n < 3; n++) {
          double en = exp(-P.beta[n][a][b] * r);
          unsafeAtomicAdd(th + 3 + n, -bb * f * en);
          unsafeAtomicAdd(th + 6 + n, bb * f * P.B[n][a][b] * r * en);
        }
      }
      add_energy(A, eatom, p.i, eb);
      D3 v = rd * (cr / r);
      atomic_add3(gc, p.i, v);
      atomic_add3(gc, p.j, -v);
      vw.push(rd * su, v);  // FP64 cm and erg/cm, in the kernel's energy unit
    }
    __syncthreads();
    gc.drain_if_crowded();
  }
  gc.flush();
  vw.flush_wave();
}

// REBO bonds with LDS-staged shells: one wavefront per tile of 64
// consecutive bonds of the bond array (consecutive centres i, since the
// array is built in atom order).  The wave first stages the shells of the
// tile's atoms (the i and j of its bonds) in LDS -- coordination sums and
// bond records, each record with the partner's index, type and total
// coordination -- then each lane evaluates one bond as k_rebo_bonds does,
// reading every shell of i and j (and of a neighbour k in the Nconj
// adjoint, when k is in the tile) from LDS.  FP32 and FP64 forms, no
// parameter gradients and no virial (those stay with k_rebo_bonds).
constexpr int kShellCache = 512;  // gradient-cache slots (LDS left by the shells)
template <class R, int OCC>
__global__ void __launch_bounds__(256, OCC) k_rebo_shell(Atoms A, const TablesT<R>* __restrict__ T,
                                                   const Tables* __restrict__ T64, Box bx,
                                                   const int2* __restrict__ bonds, const int* __restrict__ nbonds,
                                                   double* __restrict__ grad, double* __restrict__ eatom,
                                                   unsigned* err) {
  using GC = CacheOf<R, kShellCache>;
  __shared__ int s_key[GC::kSlots];
  __shared__ typename GC::val_t s_val[GC::kVals];
  __shared__ int s_fill;
  __shared__ int sh_key[4][kShHash];
  __shared__ short sh_slot[4][kShHash];
  __shared__ int sh_atom[4][kShAtoms];
  __shared__ int sh_nb[4][kShAtoms];
  __shared__ int sh_off[4][kShAtoms + 1];
  __shared__ double sh_N[4][kShAtoms][3];
  __shared__ ShellRec<R> sh_rec[4][kShRecs];
  __shared__ int sh_cnt[4];
  GC gc{s_key, s_val, &s_fill, grad, A.gfix};
  gc.clear();
  stage_rebo<R>(T);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double gs = T->eunit / T->lunit;
  const double su = 1.0 / T->eunit;  // erg -> the table's energy unit (adjoint seed)
  const BondRec<R>* B = brec<R>(A);
  Virial<false> vw;
  const LdsShell<R> lsh{A, (const MDG_LDS int*)sh_key[w], (const MDG_LDS short*)sh_slot[w],
                        (const MDG_LDS int*)sh_nb[w], (const MDG_LDS int*)sh_off[w],
                        (const MDG_LDS double*)&sh_N[w][0][0], (const MDG_LDS ShellRec<R>*)sh_rec[w]};
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  int nbt = min(*nbonds, A.bond_cap);
  int b0, b1;
  block_share(nbt, b0, b1, A.wg_map);
  for (int base = b0; base < b1; base += 256) {
    const int t = base + threadIdx.x;
    const ReboParamsT<double>& P = T64->rebo;
    int2 bd = (t < b1) ? bonds[t] : make_int2(0, 0);
    // ---- stage the shells of this wave's tile
    for (int q = lane; q < kShHash; q += 64) {
      sh_key[w][q] = -1;
      sh_slot[w][q] = -1;
    }
    if (lane == 0) sh_cnt[w] = 0;
    wave_sync();
    if (t < b1) {
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int a = e ? bd.y : bd.x;
        int h = shell_hash(a);
#pragma unroll 1
        for (int p = 0; p < kShHash; p++) {
          const int old = atomicCAS(&sh_key[w][h], -1, a);
          if (old == -1) {
            const int sl = atomicAdd(&sh_cnt[w], 1);
            sh_slot[w][h] = (short)(sl < kShAtoms ? sl : -1);
            if (sl < kShAtoms) sh_atom[w][sl] = a;
            break;
          }
          if (old == a) break;
          h = (h + 1) & (kShHash - 1);
        }
      }
    }
    wave_sync();
    const int na = min(sh_cnt[w], kShAtoms);
    // each staged atom's bond count and coordination; record offsets by a
    // prefix over the slots (an atom whose records do not fit stays global)
    int run = 0;
    for (int s0 = 0; s0 < na; s0 += 64) {
      const int sl = s0 + lane;
      int nbv = 0;
      if (sl < na) {
        const int a = sh_atom[w][sl];
        nbv = A.bcnt[a];
        sh_N[w][sl][0] = A.Nt[a];
        sh_N[w][sl][1] = A.NH[a];
        sh_N[w][sl][2] = A.NC[a];
      }
      int incl = nbv;
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      const int start = run + incl - nbv;
      const bool fits = start + nbv <= kShRecs;
      if (sl < na) {
        sh_nb[w][sl] = fits ? nbv : -1;
        sh_off[w][sl] = start;
      }
      run += __shfl(incl, 63, 64);
    }
    wave_sync();
    const int nrec = min(run, kShRecs);
    for (int r = lane; r < nrec; r += 64) {
      int lo = 0, hi = na - 1;  // last slot whose first record is <= r
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sh_off[w][mid] <= r) lo = mid;
        else hi = mid - 1;
      }
      const int tt = r - sh_off[w][lo];
      if (sh_nb[w][lo] < 0 || tt >= sh_nb[w][lo]) continue;  // records of an atom that did not fit
      const int a = sh_atom[w][lo];
      const int o = a * A.maxb + tt;
      const BondRec<R> q = B[o];
      const int k = A.bj[o];
      ShellRec<R> sr;
      sr.x = q.x;
      sr.y = q.y;
      sr.z = q.z;
      sr.r = q.r;
      sr.f = q.f;
      sr.df = q.df;
      sr.Nk = R(A.Nt[k]);
      sr.kt = ((unsigned)k << 2) | ((unsigned)A.type[k] & 3u);
      sh_rec[w][r] = sr;
    }
    wave_sync();
    // ---- the bond (as k_rebo_bonds, shells from LDS)
    BPair<R> p;
    p.i = bd.x;
    p.j = bd.y;
    p.ti = A.type[p.i];
    p.tj = A.type[p.j];
    D3 rd = depos(A.p, p.i, p.j, bx);
    p.rv = to_local<R>(rd);
    p.r = sqrt(norm2(p.rv));
    int a = p.ti, b = p.tj;
    double r = sqrt(norm2(rd));
    double f, df;
    cutoff<double, std::is_same<R, float>::value>(r, P.R[a][b][0], P.R[a][b][1], f, df);
    if (t >= b1) f = 0.0;  // past the share: no work this pass
    unsigned* const er = sys_err(err, A.sys[p.i]);
    {
      R fr, dfr;
      cutoff<R>(p.r, T->rebo.R[a][b][0], T->rebo.R[a][b][1], fr, dfr);
      p.f = fr;
    }
    p.reff = p.r;
    if (f != 0.0) {
      double cr = 0.0;
      double t1 = P.A[a][b] * exp(-P.alpha[a][b] * r);
      double val = (1.0 + P.Q[a][b] / r) * t1;
      double der = -t1 * (P.Q[a][b] + P.alpha[a][b] * r * r + P.alpha[a][b] * r * P.Q[a][b]) / (r * r);
      double VR = f * val;
      cr += der * f + val * df;
      double u1 = P.B[0][a][b] * exp(-P.beta[0][a][b] * r);
      double u2 = P.B[1][a][b] * exp(-P.beta[1][a][b] * r);
      double u3 = P.B[2][a][b] * exp(-P.beta[2][a][b] * r);
      double vsum = u1 + u2 + u3;
      double VA = -f * vsum;
      double eb = VR;
      if (VA != 0.0) {
        BOState st;
        bo_forward<R>(lsh, T, p, st, er);
        const double bb = st.b;
        bo_backward<R>(lsh, T, p, st, VA * su, gc, gs, er, vw);
        double vder = (-P.beta[0][a][b]) * u1 + (-P.beta[1][a][b]) * u2 + (-P.beta[2][a][b]) * u3;
        cr += (-vder * f * bb) + (-vsum * bb) * df;
        eb += bb * VA;
      }
      add_energy(A, eatom, p.i, eb);
      D3 v = rd * (cr / r);
      atomic_add3(gc, p.i, v);
      atomic_add3(gc, p.j, -v);
    }
    __syncthreads();
    gc.drain_if_crowded();
  }
  gc.flush();
}

// ------------------------------------------------------------ C_ij paths
// C_ij = 1 - w_ij, w_ij the best product of REBO cutoff values over paths of
// at most three bonds from i to j (AIREBO::fill_Cij / Cij, AIREBO_LJ.cxx:
// 289-408).  The reference keeps, per target, the first path met in its loop
// order (i->k, then k->m, then m->q) unless a later one has a larger weight,
// or an equal weight and fewer bonds (:311,331,352).  That is: the winner is
// the path of largest weight, then fewest bonds, then earliest in loop order.
// A path is named by key = bonds<<24 | s1<<16 | s2<<8 | s3 (its bond slots
// in the per-step bond lists, in the reference's loop order), so ties on
// weight go to the smaller key.

// A C_ij path's bonds read from the per-step bond lists (r, f, df and the
// minimum-image vector as the coordination pass computed them), loaded
// before any of the caller's scatters so that no gather waits behind them.
template <class R>
struct PathBonds {
  int hops;
  int nodes[4];
  R f[3], df[3], r[3];
  V3<R> rv[3];
};

template <class R>
__device__ __forceinline__ PathBonds<R> load_path(const Atoms& A, int i, int key) {
  const BondRec<R>* B = brec<R>(A);
  PathBonds<R> pb;
  pb.hops = key >> 24;
  pb.nodes[0] = i;
#pragma unroll
  for (int q = 0; q < 3; q++) {  // fixed trip count: the arrays stay in registers
    if (q < pb.hops) {
      int o = pb.nodes[q] * A.maxb + ((key >> (16 - 8 * q)) & 0xff);
      BondRec<R> b = B[o];
      pb.nodes[q + 1] = A.bj[o];
      pb.f[q] = b.f;
      pb.df[q] = b.df;
      pb.r[q] = b.r;
      pb.rv[q] = v3<R>(b.x, b.y, b.z);
    } else {
      pb.nodes[q + 1] = -1;
      pb.f[q] = R(1);
      pb.df[q] = R(0);
      pb.r[q] = R(1);
      pb.rv[q] = v3<R>(0, 0, 0);
    }
  }
  return pb;
}

// Path weight, product in the reference's order ((f1*f2)*f3).
template <class R>
__device__ __forceinline__ R path_weight(const PathBonds<R>& pb) {
  R w = pb.f[0];
#pragma unroll
  for (int q = 1; q < 3; q++)
    if (q < pb.hops) w = w * pb.f[q];
  return w;
}

// Gradient of C_ij = 1 - prod f over the path's bonds, seeded with V.
template <class R, class G, class Wv>
__device__ __forceinline__ void cij_backward(const PathBonds<R>& pb, R V, const G& grad, double gs, Wv& vw) {
#pragma unroll
  for (int q = 0; q < 3; q++) {
    if (q >= pb.hops || pb.df[q] == R(0)) continue;
    R mm = R(1);
#pragma unroll
    for (int u = 0; u < 3; u++)
      if (u != q && u < pb.hops) mm *= pb.f[u];
    R c = pb.df[q] * (-mm * V);
    V3<R> v = pb.rv[q] * (c / pb.r[q]);
    vw.push(pb.rv[q], v);
    scatter3<R>(grad, pb.nodes[q], v, gs);
    scatter3<R>(grad, pb.nodes[q + 1], -v, gs);
  }
}

// ---------------------------------------------------------------- LJ
template <class R>
__device__ __forceinline__ R Sfun(R a, R mn, R mx) {
  if (a < mn) return R(1);
  if (a > mx) return R(0);
  R t = (a - mn) / (mx - mn);
  return R(1) - t * t * (R(3) - R(2) * t);
}
template <class R>
__device__ __forceinline__ R dSfun(R a, R mn, R mx) {
  if (a < mn || a > mx) return R(0);
  R t = (a - mn) / (mx - mn);
  return (R(-2) * t * (R(3) - R(2) * t) + R(2) * t * t) / (mx - mn);
}

constexpr int kPathCap = 128;  // C_ij path records per atom (FP64 form)
constexpr int kHash = 64;      // C_ij hash slots per atom (power of two)
constexpr int kPairCap = 128;  // LJ pairs (j > i inside the LJ cutoff) per atom
constexpr int kL2 = 32;        // 2-bond path prefixes per atom

// Path weights w > 0 compare like their bit patterns.
template <class R>
__device__ __forceinline__ unsigned long long wbits(R w) {
  if constexpr (std::is_same<R, double>::value) return (unsigned long long)__double_as_longlong(w);
  else return (unsigned long long)__float_as_uint(w);
}
template <class R>
__device__ __forceinline__ R wfrom(unsigned long long b) {
  if constexpr (std::is_same<R, double>::value) return __longlong_as_double((long long)b);
  else return __uint_as_float((unsigned)b);
}
__device__ __forceinline__ int hslot(int t) { return (int)(((unsigned)t * 2654435761u) >> 26); }

// Sums and inclusive scans within a group of LW consecutive lanes.
template <int LW, class T>
__device__ __forceinline__ T group_sum(T v) {
  for (int o = LW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LW);
  return v;
}
template <int LW>
__device__ __forceinline__ int group_scan(int x, int sl) {
  for (int o = 1; o < LW; o <<= 1) {
    int y = __shfl_up(x, o, LW);
    if (sl >= o) x += y;
  }
  return x;
}

// LJ pair rows: for each atom i (C or H) of a system whose pflags is set,
// the j > i entries of i's long-list row (C or H) with minimum-image distance
// <= sqrt(R2sq) + margin, kept in row order.  Between two builds no atom of
// the system moved more than the margin in sum with any other, so every pair
// inside the LJ cutoff now was inside cutoff + margin then: the LJ pass finds
// exactly the pairs the full row would give it, in the same order.  One group
// of 16 lanes per atom, four chunks of loads in flight per round.
template <class R>
__global__ void __launch_bounds__(256) k_pair_rows(Atoms A, const TablesT<R>* __restrict__ T, Box bx,
                                                   const int* __restrict__ pflag, const int* __restrict__ pflags,
                                                   double margin, long long* __restrict__ rebuilt) {
  if (!*pflag) return;
  constexpr int LW = 16, U = 4, G = 256 / LW;
  const LJParamsT<R>& L = T->lj;
  const int lane = threadIdx.x & 63;
  const int sl = lane % LW;
  const int gid = threadIdx.x / LW;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mine = ((1ull << LW) - 1ull) << (lane / LW * LW);
  long long rows = 0;
  for (int i = blockIdx.x * G + gid; i < A.n; i += gridDim.x * G) {
    if (!pflags[A.sys[i]]) continue;  // group-uniform
    if (sl == 0) {
      A.disp_p[3 * i] = 0.0;
      A.disp_p[3 * i + 1] = 0.0;
      A.disp_p[3 * i + 2] = 0.0;
      rows++;
    }
    int ti = A.type[i];
    int m = 0;
    if (ti <= T_H) {
      double lim2[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        double rc = sqrt(double(L.R2sq[ti][t])) * T->lunit + margin;
        lim2[t] = rc * rc;
      }
      double4 pi4 = A.p[i];
      const Frame fr = frame_of(bx, i);
      int cnt = A.nll_cnt[i];
      const int* list = A.nll + (size_t)i * A.cap_l;
      int* out = A.pl + (size_t)i * A.cap_p;
      for (int base = A.nll_split[i]; base < cnt; base += U * LW) {
        int jj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          int q = base + u * LW + sl;
          jj[u] = (q < cnt) ? list[q] : -1;
        }
        double4 pj[U];
#pragma unroll
        for (int u = 0; u < U; u++)
          pj[u] = (jj[u] > i) ? A.p[jj[u]] : make_double4(0.0, 0.0, 0.0, double(T_OTHER));
#pragma unroll
        for (int u = 0; u < U; u++) {
          int tj = (int)pj[u].w;
          bool ok = false;
          if (jj[u] > i && (unsigned)tj <= (unsigned)T_H) {
            const D3 d = disp(bx, fr, i, jj[u], pi4, pj[u]);
            ok = d.x * d.x + d.y * d.y + d.z * d.z <= lim2[tj];
          }
          unsigned long long b = __ballot(ok) & mine;
          if (ok) {
            int slot = m + __popcll(b & below);
            if (slot < A.cap_p) out[slot] = jj[u];
          }
          m += __popcll(b);
        }
      }
    }
    if (sl == 0) A.pl_cnt[i] = (m > A.cap_p) ? -1 : m;  // -1: the LJ pass sweeps the long row
  }
  count_add(rebuilt, rows);
}

// AIREBO::operator() (AIREBO_LJ.cxx:34-77).  A group of LW lanes handles
// one atom i, so a wavefront keeps 64/LW atoms (and their chains of
// dependent gathers) in flight; LW is a tuning choice.
// Phase 0: the group sweeps the j > i part of i's long list (the row is
// sorted; its split point is stored at list build), two LW-entry chunks per
// round, and ballot-compacts the LJ pairs (inside the LJ cutoff) into LDS.
// Phase 1: every path of <= 3 bonds from i to a target j > i gets a fixed
// record index (1-bond paths, then 2-bond, then 3-bond, from prefix sums over
// the bond graph) with key = bonds<<24 | s1<<16 | s2<<8 | s3 (the reference's
// loop order), and is merged into an LDS hash keyed by target: the largest
// weight, then the smallest key -- the reference's choice whatever order the
// lanes run in.  FP32 weights and the key share one 64-bit atomic max;
// FP64 weights take an atomic max, then an atomic min on the keys of the
// paths of that weight.
// Phase 2: the group takes the compacted pairs, looks C_ij up in the hash and
// either finishes the pair (S_r = 0) or queues it for the b* kernel; all of
// a pair's gathers come before its scatters (a load issued after the
// fire-and-forget scatter atomics waits for them).
// All LDS arrays are per group, so the phases are ordered by wavefront-scope
// fences (a wave's LDS operations complete in order).
template <class R, bool TH, int OCC, int LW, bool VIR>
__global__ void __launch_bounds__(256, OCC) k_lj_wave(Atoms A, const TablesT<R>* __restrict__ T, Box bx,
                                                     double* __restrict__ grad, double* __restrict__ eatom,
                                                     int4* __restrict__ queue, int* __restrict__ qcount, int qsub,
                                                     long long* __restrict__ npairs, unsigned* err,
                                                     double* __restrict__ dth, long long* __restrict__ vir) {
  constexpr int G = 4 * (64 / LW);  // atoms in flight per workgroup
  // this workgroup's b* sub-queue
  int* const qc = qcount + (blockIdx.x & (kQueues - 1)) * kQStride;
  int4* const qb = queue + (size_t)(blockIdx.x & (kQueues - 1)) * qsub;
  constexpr bool PACK = !std::is_same<R, double>::value;
  constexpr int RC = PACK ? 1 : kPathCap;
  __shared__ int s_pair[G][kPairCap];
  __shared__ int s_ht[G][kHash];                 // target of a slot, -1 when empty
  __shared__ unsigned long long s_hw[G][kHash];  // best weight (bits) [| ~key in the FP32 form]
  __shared__ int s_hk[G][PACK ? 1 : kHash];      // smallest key among the best-weight paths (FP64)
  __shared__ int s_off[G][17];                   // 2-bond prefixes of each first bond
  __shared__ int s_m[G][kL2];                    // 2-bond prefix: end atom, weight, key bits
  __shared__ R s_w2[G][kL2];
  __shared__ int s_k2[G][kL2];
  __shared__ int s_o3[G][kL2 + 1];               // first 3-bond record of each prefix
  __shared__ short s_rs[G][RC];                  // FP64 records: hash slot, key, weight
  __shared__ int s_rk[G][RC];
  __shared__ R s_rw[G][RC];
  const LJParamsT<R>& L = T->lj;
  const BondRec<R>* B = brec<R>(A);
  const double gs = T->eunit / T->lunit;
  const int lane = threadIdx.x & 63;
  const int sl = lane % LW;
  const int gid = (threadIdx.x >> 6) * (64 / LW) + lane / LW;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mine = (LW == 64) ? ~0ull : (((1ull << LW) - 1ull) << (lane / LW * LW));
  constexpr int CS = (OCC >= 4) ? kPairCache / 2 : kPairCache;  // 4 workgroups per CU must fit the LDS
  using GC = CacheOf<R, CS>;
  __shared__ int s_key[GC::kSlots];
  __shared__ typename GC::val_t s_val[GC::kVals];
  __shared__ int s_fill;
  GC gc{s_key, s_val, &s_fill, grad, A.gfix};
  gc.clear();
  // the pair parameters of the four C/H type pairs, read from LDS per pair
  struct PairP {
    R rlj0, rlj1, rebo_r1, r0, r1, sigma, zeta;
    float rcmax, b0;
  };
  __shared__ PairP s_pp[4];
  if (threadIdx.x < 4) {
    const int a = threadIdx.x >> 1, b = threadIdx.x & 1;
    s_pp[threadIdx.x] = PairP{L.RLJ[a][b][0], L.RLJ[a][b][1], T->rebo.R[a][b][1], L.R[a][b][0], L.R[a][b][1],
                              L.sigma[a][b],   L.zeta[a][b],   float(T->rcmax[a][b]), float(L.b[a][b][0])};
  }
  __syncthreads();
  long long np = 0;
  Virial<VIR> vw(vir, T->eunit);
  // one atom per group (returns are group-uniform)
  auto atom = [&](const int i) {
    // the loads that need i alone go out together, then phase 1's bond row
    // (lane s holds bond s: partner and weight) beside phase 0's list
    const int ti = A.type[i];
    int cnt = A.pl_cnt[i];
    const double4 pi4 = A.p[i];
    const int nb1 = A.bcnt[i];
    const int bjs = (sl < nb1) ? A.bj[i * A.maxb + sl] : -1;  // nb1 <= maxb = 16 <= LW
    const R bfs = (sl < nb1) ? B[i * A.maxb + sl].f : R(0);
    if (ti > T_H) return;
    if constexpr (VIR) vw.at(A.sys[i]);
    unsigned* const er = sys_err(err, A.sys[i]);
    const Frame fr = frame_of(bx, i);
    // phase 0: compact the LJ pairs
    // the pair row, or the j > i part of the long row when the pair row overflowed
    const int* list = A.pl + (size_t)i * A.cap_p;
    int first = 0;
    if (cnt < 0) {
      cnt = A.nll_cnt[i];
      list = A.nll + (size_t)i * A.cap_l;
      first = A.nll_split[i];
    }
    int npair = 0;
    const double cut2 = T->lunit * T->lunit;
    // pairs that may have S_r != 0 (r < RLJ1, with a margin over the FP32
    // distance the pair evaluation takes) are marked: only they need the b*
    // bounds of phase 2
    double rlj2[2], lim2[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      rlj2[t] = double(L.RLJ[ti][t][1]) * double(L.RLJ[ti][t][1]) * cut2 * 1.002;
      lim2[t] = double(L.R2sq[ti][t]) * cut2;
    }
    // U chunks of LW entries per round: every list load of the round is in
    // flight before the position gathers, and those before the compaction
    constexpr int U = 4;
    for (int base = first; base < cnt; base += U * LW) {
      int jj[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        int q = base + u * LW + sl;
        jj[u] = (q < cnt) ? list[q] : -1;
      }
      double4 pj[U];
#pragma unroll
      for (int u = 0; u < U; u++) pj[u] = (jj[u] > i) ? A.p[jj[u]] : make_double4(0.0, 0.0, 0.0, double(T_OTHER));
#pragma unroll
      for (int u = 0; u < U; u++) {
        int tj = (int)pj[u].w;  // the packed position carries the type
        bool ok = false, close = false;
        if (jj[u] > i && (unsigned)tj <= (unsigned)T_H) {
          const D3 d = disp(bx, fr, i, jj[u], pi4, pj[u]);
          double d2 = d.x * d.x + d.y * d.y + d.z * d.z;
          ok = !(d2 > lim2[tj]);
          close = d2 < rlj2[tj];
        }
        unsigned long long m = __ballot(ok) & mine;
        if (ok) {
          int slot = npair + __popcll(m & below);
          if (slot < kPairCap)  // partner, its type (bit 30), the mark (bit 31)
            s_pair[gid][slot] = (int)((unsigned)jj[u] | ((unsigned)tj << 30) | (close ? 0x80000000u : 0u));
        }
        npair += __popcll(m);
      }
    }
    if (npair > kPairCap) {
      if (sl == 0) raise(er, ERR_PATH_OVERFLOW);
      npair = kPairCap;
    }
    if (npair == 0) return;
    if (sl == 0) np += npair;
    // phase 1: paths into the hash
    for (int q = sl; q < kHash; q += LW) {
      s_ht[gid][q] = -1;
      s_hw[gid][q] = 0ull;
      if constexpr (!PACK) s_hk[gid][q] = 0x7fffffff;
    }
    int c2 = 0;
    if (sl < nb1) c2 = A.bcnt[bjs];
    int incl = group_scan<LW>(c2, sl);
    if (sl < 16) s_off[gid][sl + 1] = incl;
    if (sl == 0) {
      s_off[gid][0] = 0;
      s_o3[gid][0] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int tot2 = s_off[gid][nb1];
    if (tot2 > kL2) {
      if (sl == 0) raise(er, ERR_PATH_OVERFLOW);
      tot2 = kL2;
    }
    // 2-bond prefixes (s1, s2); the first bond's partner and weight come from
    // lane s1 (tot2 is the same for the whole group: all its lanes shuffle)
    for (int c0 = 0; c0 < tot2; c0 += LW) {
      const int c = c0 + sl;
      int s1 = 0;
      while (s1 + 1 < nb1 && s_off[gid][s1 + 1] <= c) s1++;
      const int k = __shfl(bjs, s1, LW);
      const R f1 = __shfl(bfs, s1, LW);
      if (c < tot2) {
        int s2 = c - s_off[gid][s1];
        int m = A.bj[k * A.maxb + s2];
        s_m[gid][c] = m;
        s_w2[gid][c] = f1 * B[k * A.maxb + s2].f;
        s_k2[gid][c] = (s1 << 16) | (s2 << 8);
        s_o3[gid][c + 1] = A.bcnt[m];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int base = 0, run = 0; base < tot2; base += LW) {  // s_o3[c]: first 3-bond record of prefix c
      int c = base + sl;
      int x = (c < tot2) ? s_o3[gid][c + 1] : 0;
      x = group_scan<LW>(x, sl);
      if (c < tot2) s_o3[gid][c + 1] = run + x;
      run += __shfl(x, LW - 1, LW);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int nrec = nb1 + tot2 + s_o3[gid][tot2];
    if (!PACK && nrec > kPathCap) {
      if (sl == 0) raise(er, ERR_PATH_OVERFLOW);
      nrec = kPathCap;
    }
    for (int r = sl; r < nrec; r += LW) {
      int target, key;
      R w;
      if (r < nb1) {  // r < 16 <= LW: this lane's own bond
        target = bjs;
        w = bfs;
        key = (1 << 24) | (r << 16);
      } else if (r < nb1 + tot2) {
        int c = r - nb1;
        target = s_m[gid][c];
        w = s_w2[gid][c];
        key = (2 << 24) | s_k2[gid][c];
      } else {
        int d = r - nb1 - tot2;
        int lo = 0, hi = tot2 - 1;  // last prefix whose first record is <= d
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (s_o3[gid][mid] <= d) lo = mid;
          else hi = mid - 1;
        }
        int s3 = d - s_o3[gid][lo];
        int m = s_m[gid][lo];
        target = A.bj[m * A.maxb + s3];
        w = s_w2[gid][lo] * B[m * A.maxb + s3].f;
        key = (3 << 24) | s_k2[gid][lo] | s3;
      }
      int h = -1;
      if (target > i && w > R(0)) {  // only targets j > i are LJ pairs of this atom
        h = hslot(target);
        for (int probes = 0;; probes++) {
          if (probes == kHash) {
            raise(er, ERR_PATH_OVERFLOW);
            h = -1;
            break;
          }
          int t = atomicCAS(&s_ht[gid][h], -1, target);
          if (t == -1 || t == target) break;
          h = (h + 1) & (kHash - 1);
        }
        if (h >= 0) {
          if constexpr (PACK) atomicMax(&s_hw[gid][h], (wbits(w) << 32) | (unsigned long long)(~(unsigned)key));
          else atomicMax(&s_hw[gid][h], wbits(w));
        }
      }
      if constexpr (!PACK) {
        s_rs[gid][r] = (short)h;
        s_rk[gid][r] = key;
        s_rw[gid][r] = w;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if constexpr (!PACK) {
      for (int r = sl; r < nrec; r += LW) {
        int h = s_rs[gid][r];
        if (h >= 0 && wbits(s_rw[gid][r]) == s_hw[gid][h]) atomicMin(&s_hk[gid][h], s_rk[gid][r]);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    // phase 2: pairs
    double e = 0.0;
    double th_e[2] = {0.0, 0.0}, th_s[2] = {0.0, 0.0};  // dE/d(epsilon, sigma) by tj
    D3 gi = d3(0, 0, 0);  // FP64 in both forms: the lane width only reorders FP64 sums
    // one pair (j, its packed position); the loads of a lane's next two
    // pairs are issued together below
    auto pair = [&](const int j, const double4 pj, const float bsum) {
      const int tj = (int)pj.w;  // phase 0 kept C/H partners only
      if ((unsigned)tj > (unsigned)T_H) {  // a stale packed copy: never index the tables with it
        raise(er, ERR_TYPE_RANGE);
        return;
      }
      const PairP& pp = s_pp[2 * ti + tj];
      D3 rd = disp(bx, fr, i, j, pi4, pj);
      R w = R(0);
      int bkey = 0x7fffffff;
      int h = hslot(j);
      for (int probes = 0; probes < kHash; probes++) {
        int t = s_ht[gid][h];
        if (t == j) {
          unsigned long long v = s_hw[gid][h];
          if constexpr (PACK) {
            w = wfrom<R>(v >> 32);
            bkey = (int)(~(unsigned)(v & 0xffffffffull));
          } else {
            w = wfrom<R>(v);
            bkey = s_hk[gid][h];
          }
          break;
        }
        if (t == -1) break;
        h = (h + 1) & (kHash - 1);
      }
      R C = R(1) - w;
      if (!(C > R(0))) return;
      int pkey = (w > R(0)) ? bkey : 0;  // C_ij path (0: none)
      V3<R> rv = to_local<R>(rd);
      R r = sqrt(norm2(rv));
      R Str = Sfun<R>(r, pp.rlj0, pp.rlj1);
      // b* < b_min by the per-atom bounds (a pair that is no REBO bond:
      // neither atom's sum holds the other, and pi_rc <= rcmax, pi_dh <= 0):
      // then S_b = 1 and the pair is one of S_r = 0's form, finished here
      if (Str != R(0) && A.pmax && r > pp.rebo_r1 &&
          0.5f * bsum + pp.rcmax + 1e-5f < pp.b0)
        Str = R(0);
      if (Str != R(0)) {  // b* needed: one queue atomic per wave
        unsigned long long m = __ballot(1);
        int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(qc, __popcll(m));
        base = __shfl(base, leader, 64);
        int slot = base + __popcll(m & below);
        if (slot < qsub) qb[slot] = make_int4(i, j, pkey, 0);
        else raise(err, ERR_QUEUE_OVERFLOW);
        return;
      }
      PathBonds<R> pb;
      pb.hops = 0;
      if (pkey) pb = load_path<R>(A, i, pkey);
      R f, df;
      cutoff<R>(r, pp.r0, pp.r1, f, df);
      if (f == R(0)) return;
      R s = pp.sigma / r;
      R s6 = s;
      for (int u = 2; u <= 6; u++) s6 *= s;
      R s12 = s6 * s6;
      R z = pp.zeta;
      R val = R(4) * z * (s12 - s6);
      R der = R(4) * z * (R(-12) * s12 / r + R(6) * s6 / r);
      R LJ = f * val;
      R V = R(1) * C;  // VLJ seeded with SrSb*C, SrSb = 1
      R cr = der * f * V + val * V * df;
      V3<R> v = rv * (cr / r);
      gi = gi + d3(double(v.x), double(v.y), double(v.z));
      vw.push(rv, v);
      scatter3<R>(gc, j, -v, gs);
      if (pkey) cij_backward<R>(pb, R(1) * LJ, gc, gs, vw);
      e += to_erg<R>(double(R(1) * C * LJ), T->eunit);
      if constexpr (TH) {  // S_r = 0: E = C f 4 eps ((sigma/r)^12 - (sigma/r)^6)
        double de = double(C) * double(f) * 4.0 * double(s12 - s6);
        double ds = double(C) * double(f) * 4.0 * double(z) * double(R(12) * s12 - R(6) * s6) /
                    double(pp.sigma);
        if (tj == T_C) {
          th_e[0] += de;
          th_s[0] += ds;
        } else {
          th_e[1] += de;
          th_s[1] += ds;
        }
      }
    };
    // a lane's next two pairs: positions and, for marked pairs, the per-atom
    // b* bounds (pmax[i][tj] + pmax[j][ti]) gathered together; an unmarked
    // pair's bound sum stays 1e30 (nothing proven)
    for (int q = sl; q < npair; q += 2 * LW) {
      const int q2 = q + LW;
      const int sa = s_pair[gid][q];
      const int sb = (q2 < npair) ? s_pair[gid][q2] : sa;
      const int ja = sa & 0x3fffffff, jb = sb & 0x3fffffff;
      const double4 pa = A.p[ja], pb4 = A.p[jb];
      float ba = 1e30f, bb = 1e30f;
      if (A.pmax) {
        if (sa < 0) ba = A.pmax[2 * i + ((sa >> 30) & 1)] + A.pmax[2 * ja + ti];
        if (sb < 0) bb = A.pmax[2 * i + ((sb >> 30) & 1)] + A.pmax[2 * jb + ti];
      }
      pair(ja, pa, ba);
      if (q2 < npair) pair(jb, pb4, bb);
    }
    double gx = group_sum<LW>(gi.x), gy = group_sum<LW>(gi.y), gz = group_sum<LW>(gi.z);
    e = group_sum<LW>(e);
    if (sl == 0) {
      if (gx != 0.0 || gy != 0.0 || gz != 0.0) {
        if constexpr (std::is_same<R, double>::value) atomic_add3(gc, i, d3(gx, gy, gz));
        else atomic_add3(gc, i, d3(gx * gs, gy * gs, gz * gs));
      }
      add_energy(A, eatom, i, e);
    }
    if constexpr (TH) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        double de = group_sum<LW>(th_e[q]), ds = group_sum<LW>(th_s[q]);
        if (sl == 0 && (de != 0.0 || ds != 0.0)) {
          double* th = theta_of(dth, A, i) + 27 + 2 * (ti + q);
          unsafeAtomicAdd(th, de);
          unsafeAtomicAdd(th + 1, ds);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  // a contiguous share of the atoms per workgroup, G atoms per pass
  int a0, a1;
  block_share(A.n, a0, a1, A.wg_map, G);
    for (int base = a0; base < a1; base += G) {
      if (base + gid < a1) atom(base + gid);
      __syncthreads();
      gc.drain_if_crowded();
    }
  gc.flush();
  vw.flush_wave();
  count_add(npairs, np);
}

// LJ pairs that need b*_ij (S_r != 0): the whole StrStb / VLJ / Cij chain
// (AIREBO_LJ.cxx:79-202).
//
// b* < b_min needs no more than the sigma-pi sums: pi_dh = T_ij * sum
// (1 - cos^2) f' f' <= 0 (T_ij's nodes are <= 0 with zero slopes, so its
// interpolant stays between them) and pi_rc <= rcmax, so
//   b* <= bub = (D_ij^-1/2 + D_ji^-1/2)/2 + rcmax
// (taken with 1/sqrt and a 1e-12 slack that covers every rounding
// difference to the full evaluation's pow).  When bub < b_min, S(b*) = 1
// and dS(b*) = 0 exactly, as the full evaluation would find, and neither
// Nconj, pi_rc, pi_dh nor the bond-order adjoint is needed.
// PASS 0: one pass, the full evaluation inline where the bound fails.
// PASS 1: the bound pass; a pair whose bound fails is marked (queue entry w
//         = 1) and left to PASS 2 entirely.  It compiles without the bond
//         order's second half and adjoint, so it keeps few registers.
// PASS 2: the full evaluation of the marked pairs.
template <class R, bool TH, int OCC, int PASS, bool VIR>
__global__ void __launch_bounds__(256, OCC) k_bstar(Atoms A, const TablesT<R>* __restrict__ T, Box bx,
                                              int4* __restrict__ queue, const int* __restrict__ qcount,
                                              int qsub, double* __restrict__ grad, double* __restrict__ eatom,
                                              unsigned* err, double* __restrict__ dth, long long* __restrict__ vir,
                                              long long* __restrict__ work) {
  static_assert(kQueues == 256, "one thread per sub-queue in the prefix below (blockDim 256)");
  static_assert(PASS >= 0 && PASS <= 2 && (PASS == 0 || !TH), "parameter gradients use the one-pass form");
  const double gs = T->eunit / T->lunit;
  // s_pre[q]: first flat index of sub-queue q (exclusive prefix of the counts)
  __shared__ int s_pre[kQueues + 1];
  __shared__ int s_wsum[4];
  {
    int q = threadIdx.x;
    int c = min(qcount[q * kQStride], qsub);
    int x = c;  // inclusive scan within the wave, then across the four waves
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(x, o, 64);
      if ((q & 63) >= o) x += y;
    }
    if ((q & 63) == 63) s_wsum[q >> 6] = x;
    __syncthreads();
    for (int w = 0; w < (q >> 6); w++) x += s_wsum[w];
    s_pre[q + 1] = x;
    if (q == 0) s_pre[0] = 0;
    __syncthreads();
  }
  using GC = CacheOf<R, kBondCache>;
  __shared__ int s_key[GC::kSlots];
  __shared__ typename GC::val_t s_val[GC::kVals];
  __shared__ int s_fill;
  GC gc{s_key, s_val, &s_fill, grad, A.gfix};
  gc.clear();
  stage_rebo<R>(T);
  __syncthreads();
  const int nq = s_pre[kQueues];
  Virial<VIR> vw(vir, T->eunit);
  long long n_trip = 0, n_full = 0;  // PASS 1 work tallies
  int b0, b1;
  block_share(nq, b0, b1, A.wg_map);
  for (int base = b0; base < b1; base += 256) {
   const int t = base + threadIdx.x;
   int lo = 0;
   int4 qd = make_int4(0, 0, 0, 0);
   bool work = t < b1;
   if (work) {
    int hi = kQueues - 1;  // last sub-queue whose first index is <= t
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    qd = queue[(size_t)lo * qsub + (t - s_pre[lo])];
    if constexpr (PASS == 2) work = qd.w != 0;
   }
   if (work) {
    const LJParamsT<R>& L = T->lj;
    const ReboParamsT<R>& P = T->rebo;
    int i = qd.x, j = qd.y;
    int pkey = qd.z;  // C_ij path (0: none), bond slots as in k_lj_wave
    int ti = A.type[i], tj = A.type[j];
    unsigned* const er = sys_err(err, A.sys[i]);
    V3<R> rv = to_local<R>(depos(A.p, i, j, bx));
    R r = sqrt(norm2(rv));
    PathBonds<R> pb;
    pb.hops = 0;
    if (pkey) pb = load_path<R>(A, i, pkey);
    R w = pkey ? path_weight<R>(pb) : R(0);
    R C = R(1) - w;
    R rmin = L.RLJ[ti][tj][0], rmax = L.RLJ[ti][tj][1];
    R Str = Sfun<R>(r, rmin, rmax);
    // b* pair: REBO cutoffs at the actual geometry, r_alter = R0 in ExpTerm
    BPair<R> p;
    p.i = i;
    p.j = j;
    p.ti = ti;
    p.tj = tj;
    p.rv = rv;
    p.r = r;
    R fb, dfb;
    cutoff<R>(r, P.R[ti][tj][0], P.R[ti][tj][1], fb, dfb);
    p.f = fb;
    p.reff = P.R[ti][tj][0];
    BOState st;
    R bmin = L.b[ti][tj][0], bmax = L.b[ti][tj][1];
    bo_sigma_pi<R>(GlobalShell<R>{A}, T, p, st, er);
    bool below = false;
    if constexpr (PASS != 2) {
      const double bub = 0.5 * (1.0 / sqrt(st.D[0]) + 1.0 / sqrt(st.D[1])) + T->rcmax[ti][tj] + 1e-12;
      below = bub < double(bmin);
    }
    bool done = false;
    if constexpr (PASS == 1) {
      n_trip += A.bcnt[i] + A.bcnt[j];  // sigma-pi triplets of the bound (both sides, no partner excluded)
      if (!below) {  // left to the full pass
        queue[(size_t)lo * qsub + (t - s_pre[lo])].w = 1;
        done = true;
        n_full++;
      }
    } else {
      if (!below) bo_finish<R>(GlobalShell<R>{A}, T, p, st, er);
    }
    if (!done) {
    if constexpr (VIR) vw.at(A.sys[i]);
    R bs = below ? bmin : R(st.b);
    R Stb = below ? R(1) : Sfun<R>(bs, bmin, bmax);
    R SrSb = R(1) + (Str * Stb - Str);
    R f, df;
    cutoff<R>(r, L.R[ti][tj][0], L.R[ti][tj][1], f, df);
    R s = L.sigma[ti][tj] / r;
    R s6 = s;
    for (int u = 2; u <= 6; u++) s6 *= s;
    R s12 = s6 * s6;
    R z = L.zeta[ti][tj];
    R val = R(4) * z * (s12 - s6);
    R der = R(4) * z * (R(-12) * s12 / r + R(6) * s6 / r);
    R LJ = (f == R(0)) ? R(0) : f * val;
    R cr = R(0);
    if (f != R(0)) {
      R V = SrSb * C;
      cr += der * f * V + val * V * df;
    }
    // StrStb seeded with C*LJ
    R V2 = C * LJ;
    if (V2 != R(0)) {
      R dStr = dSfun<R>(r, rmin, rmax);
      if (dStr != R(0)) cr += -dStr * V2;
      if (dStr != R(0) && Stb != R(0)) cr += dStr * Stb * V2;
      if constexpr (PASS != 1) {
        R dStb = below ? R(0) : dSfun<R>(bs, bmin, bmax);
        if (dStb != R(0)) bo_backward<R>(GlobalShell<R>{A}, T, p, st, double(Str * dStb * V2), gc, gs, er, vw);
      }
    }
    V3<R> v = rv * (cr / r);
    scatter3<R>(gc, i, v, gs);
    scatter3<R>(gc, j, -v, gs);
    vw.push(rv, v);
    if (pkey && SrSb * LJ != R(0)) cij_backward<R>(pb, SrSb * LJ, gc, gs, vw);
    add_energy(A, eatom, i, to_erg<R>(double(SrSb * C * LJ), T->eunit));
    if constexpr (TH) {
      // S_r depends on r/sigma only: dS_r/dsigma = -(r/sigma) dS_r/dr
      double* th = theta_of(dth, A, i) + 27 + 2 * (ti + tj);
      double sg = double(L.sigma[ti][tj]);
      double fv = (f == R(0)) ? 0.0 : double(f);
      double de = double(SrSb) * double(C) * fv * 4.0 * double(s12 - s6);
      double ds = double(SrSb) * double(C) * fv * 4.0 * double(z) * double(R(12) * s12 - R(6) * s6) / sg +
                  (-(double(r) / sg) * double(dSfun<R>(r, rmin, rmax))) * double(Stb - R(1)) * double(C) *
                      double(LJ);
      unsafeAtomicAdd(th, de);
      unsafeAtomicAdd(th + 1, ds);
    }
    }
   }
    __syncthreads();
    gc.drain_if_crowded();
  }
  gc.flush();
  vw.flush_wave();
  if constexpr (PASS == 1) {
    count_add(work + 0, n_trip);
    count_add(work + 1, n_full);
  }
}

// ------------------------------------------------------------ torsion
// ETors::operator() (AIREBO_ETors.cxx:32-115) over C-C bonds.
template <class R, bool VIR>
__global__ void __launch_bounds__(256) k_etors(Atoms A, const TablesT<R>* __restrict__ T, Box bx,
                                              const int2* __restrict__ bonds, const int* __restrict__ nbonds,
                                              int compat, double* __restrict__ grad, double* __restrict__ eatom,
                                              long long* __restrict__ nquads, long long* __restrict__ vir) {
  using GC = CacheOf<R, kTorsCache>;
  __shared__ int s_key[GC::kSlots];
  __shared__ typename GC::val_t s_val[GC::kVals];
  __shared__ int s_fill;
  GC gc{s_key, s_val, &s_fill, grad, A.gfix};
  gc.clear();
  __syncthreads();
  long long nq = 0;
  const double gs = T->eunit / T->lunit;
  Virial<VIR> vw(vir, T->eunit);
  const R zt_cc = T->zetaT[T_C][T_C], zt_ch = T->zetaT[T_C][T_H], zt_hc = T->zetaT[T_H][T_C],
          zt_hh = T->zetaT[T_H][T_H];
  int nbt = min(*nbonds, A.bond_cap);
  int b0, b1;
  block_share(nbt, b0, b1, A.wg_map);
  for (int base = b0; base < b1; base += 256) {
    const int t = base + threadIdx.x;
    const ReboParamsT<R>& P = T->rebo;
    int2 bd = (t < b1) ? bonds[t] : make_int2(0, 0);
    int i = bd.x, j = bd.y;
    if (t < b1 && A.type[i] == T_C && A.type[j] == T_C) {
      if constexpr (VIR) vw.at(A.sys[i]);
      V3<R> a_ = to_local<R>(depos(A.p, i, j, bx));
      R rij = sqrt(norm2(a_));
      R wij, dwij;
      cutoff<R>(rij, P.R[0][0][0], P.R[0][0][1], wij, dwij);
      V3<R> gi = v3<R>(0, 0, 0), gj = v3<R>(0, 0, 0);
      int nbi = A.bcnt[i], nbj = A.bcnt[j];
      double eb = 0.0;
      for (int s = 0; s < nbi; s++) {
        BondRef<R> k = bond_at<R>(A, i, s);
        if (k.k == j) continue;
        R sin_ik = sqrt(norm2(cross(a_, k.rv))) / (rij * k.r);
        for (int u = 0; u < nbj; u++) {
          BondRef<R> l = bond_at<R>(A, j, u);
          if (l.k == i || l.k == k.k) continue;
          if (fabs(sin_ik) < R(0.1)) continue;
          R sin_jl = sqrt(norm2(cross(a_, l.rv))) / (rij * l.r);
          if (fabs(sin_jl) < R(0.1)) continue;
          nq++;
          int tk = A.type[k.k], tl = A.type[l.k];
          R wki = k.f, wjl = l.f;
          // zeta of (tk, tl) by selects from the four values held in registers
          R zt = (tk == T_C) ? ((tl == T_C) ? zt_cc : zt_ch) : ((tl == T_C) ? zt_hc : zt_hh);
          V3<R> b_ = k.rv, c_ = l.rv;
          V3<R> Pp = cross(-a_, b_), Q = cross(a_, c_);
          R p2 = norm2(Pp), q2 = norm2(Q);
          R e2 = sqrt(p2) * sqrt(q2);
          R cs = (e2 == R(0)) ? R(0) : dot(Pp, Q) / e2;
          R cd = cs;
          if (cd < R(-1)) cd = R(-1);
          if (cd > R(1)) cd = R(1);
          R c = -cd;
          R x = R(0.5) * (R(1) + c);
          // x^4 and x^5: the library pow in the FP64 form (as the reference),
          // products in the FP32 one (x in [0, 1]: a few ulp)
          R x4, x5;
          if constexpr (std::is_same<R, float>::value) {
            const R xx = x * x;
            x4 = xx * xx;
            x5 = x4 * x;
          } else {
            x5 = pow(x, R(5));
            x4 = pow(x, R(5) - R(1));
          }
          R vt = R(256.0 / 405.0) * zt * x5 - R(0.1) * zt;
          V3<R> gk = v3<R>(0, 0, 0), gl = v3<R>(0, 0, 0);
          if (vt != R(0)) {
            R der = R(256.0 / 405.0) * zt * x4 * R(0.5);
            if (!compat) der *= R(5);
            R V = R(-1) * der * (wki * wij * wjl);
            if (V != R(0) && e2 != R(0)) {
              V3<R> gP = Q / e2 - Pp * (cs / p2);
              V3<R> gQ = Pp / e2 - Q * (cs / q2);
              V3<R> da = cross(gP, b_) + cross(c_, gQ);
              V3<R> db = cross(a_, gP);
              V3<R> dc = cross(gQ, a_);
              gi = gi + (da + db) * V;
              gj = gj + (dc - da) * V;
              gk = gk - db * V;
              gl = gl - dc * V;
              vw.push(a_, da * V);
              vw.push(b_, db * V);
              vw.push(c_, dc * V);
            }
          }
          // cutoff-weight derivatives
          if (k.df != R(0)) {
            V3<R> v = k.rv * (k.df * (wij * wjl * vt) / k.r);
            gi = gi + v;
            gk = gk - v;
            vw.push(k.rv, v);
          }
          if (dwij != R(0)) {
            V3<R> v = a_ * (dwij * (wki * wjl * vt) / rij);
            gi = gi + v;
            gj = gj - v;
            vw.push(a_, v);
          }
          if (l.df != R(0)) {
            V3<R> v = l.rv * (l.df * (wki * wij * vt) / l.r);
            gj = gj + v;
            gl = gl - v;
            vw.push(l.rv, v);
          }
          scatter3<R>(gc, k.k, gk, gs);
          scatter3<R>(gc, l.k, gl, gs);
          eb += double(wki * wij * wjl * vt);
        }
      }
      add_energy(A, eatom, i, to_erg<R>(eb, T->eunit));
      scatter3<R>(gc, i, gi, gs);
      scatter3<R>(gc, j, gj, gs);
    }
    __syncthreads();
    gc.drain_if_crowded();
  }
  gc.flush();
  vw.flush_wave();
  count_add(nquads, nq);
}

// The FP64 form's kernels and the mixed form's are compiled in two passes
// of this file (the build defines MDG_FORCE_FP64 or MDG_FORCE_FP32; both when
// neither is given): the mixed form's with FMA contraction, the FP64 form's
// without, as the reference's x86-64 build rounds.
#if !defined(MDG_FORCE_FP64) && !defined(MDG_FORCE_FP32)
#define MDG_FORCE_FP64
#define MDG_FORCE_FP32
#endif
#define MDG_INSTANTIATE_OCC(R, TH, OCC)                                                                        \
  template __global__ void k_rebo_bonds<R, TH, OCC, false>(Atoms, const TablesT<R>*, const Tables*, Box,          \
                                                           const int2*, const int*, double*, double*, unsigned*, \
                                                           double*, long long*);                                    \
  template __global__ void k_bstar<R, TH, OCC, 0, false>(Atoms, const TablesT<R>*, Box, int4*, const int*, int,  \
                                                         double*, double*, unsigned*, double*, long long*, long long*);
#define MDG_INSTANTIATE_TH(R, TH, OCC, LW)                                                                     \
  template __global__ void k_lj_wave<R, TH, OCC, LW, false>(Atoms, const TablesT<R>*, Box, double*, double*,   \
                                                            int4*, int*, int, long long*, unsigned*, double*,   \
                                                            long long*);
#define MDG_INSTANTIATE(R)                                                                                     \
  template __global__ void k_coord<R>(Atoms, const TablesT<R>*, Box, unsigned*);                              \
  template __global__ void k_etors<R, false>(Atoms, const TablesT<R>*, Box, const int2*, const int*, int,     \
                                             double*, double*, long long*, long long*);                          \
  template __global__ void k_etors<R, true>(Atoms, const TablesT<R>*, Box, const int2*, const int*, int,      \
                                            double*, double*, long long*, long long*);                           \
  template __global__ void k_rebo_bonds<R, false, 1, true>(Atoms, const TablesT<R>*, const Tables*, Box,      \
                                                           const int2*, const int*, double*, double*,         \
                                                           unsigned*, double*, long long*);                      \
  template __global__ void k_lj_wave<R, false, 1, 16, true>(Atoms, const TablesT<R>*, Box, double*, double*,  \
                                                            int4*, int*, int, long long*, unsigned*, double*,  \
                                                            long long*);                                          \
  template __global__ void k_bstar<R, false, 1, 0, true>(Atoms, const TablesT<R>*, Box, int4*, const int*,    \
                                                         int, double*, double*, unsigned*, double*, long long*, long long*);  \
  template __global__ void k_pair_rows<R>(Atoms, const TablesT<R>*, Box, const int*, const int*, double,      \
                                          long long*);
#define MDG_INSTANTIATE_SHELL(R, OCC)                                                                        \
  template __global__ void k_rebo_shell<R, OCC>(Atoms, const TablesT<R>*, const Tables*, Box, const int2*,       \
                                                const int*, double*, double*, unsigned*);
#define MDG_INSTANTIATE_SPLIT(R, OCC)                                                                           \
  template __global__ void k_bstar<R, false, OCC, 1, false>(Atoms, const TablesT<R>*, Box, int4*, const int*,  \
                                                            int, double*, double*, unsigned*, double*,         \
                                                            long long*, long long*);                                          \
  template __global__ void k_bstar<R, false, OCC, 2, false>(Atoms, const TablesT<R>*, Box, int4*, const int*,  \
                                                            int, double*, double*, unsigned*, double*,         \
                                                            long long*, long long*);
#ifdef MDG_FORCE_FP64
MDG_INSTANTIATE(double)
MDG_INSTANTIATE_SHELL(double, 1)
MDG_INSTANTIATE_TH(double, false, 1, 64)
MDG_INSTANTIATE_TH(double, false, 3, 64)
MDG_INSTANTIATE_TH(double, false, 1, 32)
MDG_INSTANTIATE_TH(double, false, 1, 16)
MDG_INSTANTIATE_TH(double, true, 1, 16)
MDG_INSTANTIATE_OCC(double, false, 1)
MDG_INSTANTIATE_OCC(double, false, 2)
MDG_INSTANTIATE_OCC(double, true, 1)
MDG_INSTANTIATE_SPLIT(double, 1)
MDG_INSTANTIATE_SPLIT(double, 2)
#endif
#ifdef MDG_FORCE_FP32
MDG_INSTANTIATE(float)
MDG_INSTANTIATE_SHELL(float, 1)
MDG_INSTANTIATE_SHELL(float, 2)
MDG_INSTANTIATE_TH(float, false, 1, 64)
MDG_INSTANTIATE_TH(float, false, 3, 64)
MDG_INSTANTIATE_TH(float, false, 1, 32)
MDG_INSTANTIATE_TH(float, false, 1, 16)
MDG_INSTANTIATE_TH(float, false, 4, 16)
MDG_INSTANTIATE_TH(float, false, 3, 16)
MDG_INSTANTIATE_OCC(float, false, 1)
MDG_INSTANTIATE_OCC(float, false, 2)
MDG_INSTANTIATE_OCC(float, false, 3)
MDG_INSTANTIATE_OCC(float, false, 4)
MDG_INSTANTIATE_SPLIT(float, 1)
MDG_INSTANTIATE_SPLIT(float, 2)
MDG_INSTANTIATE_SPLIT(float, 3)
MDG_INSTANTIATE_SPLIT(float, 4)
#endif

}  // namespace mdg
