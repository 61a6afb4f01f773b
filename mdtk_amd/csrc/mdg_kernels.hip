// FP64 kernels of the force-and-integration path.
//
// Force scheme: the reference evaluates every energy sub-term once for its
// value and again for each derivative seed (REBO.cxx:200-814).  Here every
// bond (and every b* pair of the LJ term) runs one forward pass that keeps the
// scalar intermediates, then one adjoint pass seeded with dE/db that scatters
// V*dterm/dx into the gradient with FP64 atomics.  Geometry, cutoffs and the
// summation order of every value follow the reference; kernels are compiled
// without FMA contraction so that a*b+c rounds twice, as there.
#include "mdg_kernels.hpp"

namespace mdg {

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ void raise(unsigned* err, unsigned bit) { atomicOr(err, bit); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sum of one double per thread, one atomic per block into *acc.
__device__ void block_accumulate(double v, double* acc) {
  __shared__ double red[16];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) s += red[q];
    if (s != 0.0) unsafeAtomicAdd(acc, s);
  }
  __syncthreads();
}

struct BondRef {
  int k;
  double r, f, df;
  D3 rv;  // x_a - x_k (minimum image)
};

__device__ __forceinline__ BondRef bond_at(const Atoms& A, int a, int s) {
  int o = a * A.maxb + s;
  BondRef b;
  b.k = A.bj[o];
  b.r = A.br[o];
  b.f = A.bf[o];
  b.df = A.bdf[o];
  b.rv = d3(A.bx[o], A.by[o], A.bz[o]);
  return b;
}

// Coordination of a excluding b (REBO::Nt/NH/NC, REBO.cxx:423-493): the
// cached sum minus f_ab when r_ab < R2(a,b), clamped at 0.
__device__ __forceinline__ double coord_excl(double Nsum, bool subtract, double r, double R2, double fab,
                                             unsigned* err) {
  double N = Nsum;
  if (subtract && r < R2) N -= fab;
  if (!(N > -1e-15)) raise(err, ERR_NEG_COORD);
  if (N < 0) N = 0;
  return N;
}

// ------------------------------------------------------------ cell lists
__global__ void k_cell_assign(const int* __restrict__ flag, int n, const double* __restrict__ x,
                              const double* __restrict__ y, const double* __restrict__ z, CellGrid g,
                              int* __restrict__ cell_of, int* __restrict__ cell_cnt) {
  if (!*flag) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
  double p[3] = {x[i], y[i], z[i]};
  for (int d = 0; d < 3; d++) {
    int q = (int)floor((p[d] - g.lo[d]) / g.w[d]);
    if (q < 0) q = 0;
    if (q >= g.n[d]) q = g.n[d] - 1;
    c[d] = q;
  }
  int id = (c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
  cell_of[i] = id;
  atomicAdd(&cell_cnt[id], 1);
}

// Exclusive scan of the cell counts by one workgroup (1024 threads).
__global__ void k_cell_scan(const int* __restrict__ flag, int ncell, int* __restrict__ cnt, int* __restrict__ start) {
  if (!*flag) return;
  __shared__ int part[1024];
  int t = threadIdx.x;
  int chunk = (ncell + 1023) / 1024;
  int b = t * chunk, e = min(b + chunk, ncell);
  int s = 0;
  for (int q = b; q < e; q++) s += cnt[q];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = (t >= o) ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int q = b; q < e; q++) {
    start[q] = run;
    run += cnt[q];
    cnt[q] = 0;  // reused as fill cursor
  }
  if (t == 1023) start[ncell] = part[1023];
}

__global__ void k_cell_fill(const int* __restrict__ flag, int n, const int* __restrict__ cell_of,
                            const int* __restrict__ start, int* __restrict__ cursor, int* __restrict__ cell_atoms) {
  if (!*flag) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of[i];
  int p = atomicAdd(&cursor[c], 1);
  cell_atoms[start[c] + p] = i;
}

// Neighbour list build, one wavefront per atom (NeighbourList.cxx:31-106):
// pair (i,j) is listed iff both atoms are handled and d^2 < ((1+0.44) Rc)^2
// with d^2 = dx*dx+dy*dy+dz*dz of the minimum-image displacement.  Candidates
// from the stencil cells are compacted with a ballot/popcount prefix into
// LDS, then sorted ascending (the reference's list order) with a bitonic sort.
template <int CAP>
__global__ void __launch_bounds__(256) k_nl_build(const int* __restrict__ flag, int n, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ z,
                                                  const unsigned char* __restrict__ handled, Box bx, CellGrid g,
                                                  double range2, const int* __restrict__ cell_of,
                                                  const int* __restrict__ start, const int* __restrict__ cell_atoms,
                                                  int* __restrict__ cnt, int* __restrict__ list,
                                                  double* __restrict__ disp, unsigned* err) {
  if (!*flag) return;
  __shared__ int buf[4][CAP];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int i = blockIdx.x * 4 + w;
  if (i >= n) return;
  int* my = buf[w];
  if (lane == 0) {
    disp[3 * i] = 0.0;
    disp[3 * i + 1] = 0.0;
    disp[3 * i + 2] = 0.0;
  }
  int m = 0;
  if (handled[i]) {
    double xi = x[i], yi = y[i], zi = z[i];
    int ci = cell_of[i];
    int c2 = ci % g.n[2], c1 = (ci / g.n[2]) % g.n[1], c0 = ci / (g.n[1] * g.n[2]);
    int lo[3], hi[3], cc[3] = {c0, c1, c2};
    for (int d = 0; d < 3; d++) {
      if (g.n[d] >= 3) {
        lo[d] = -1;
        hi[d] = 1;
      } else {
        lo[d] = -cc[d];
        hi[d] = g.n[d] - 1 - cc[d];
      }
    }
    for (int a = lo[0]; a <= hi[0]; a++)
      for (int b = lo[1]; b <= hi[1]; b++)
        for (int c = lo[2]; c <= hi[2]; c++) {
          int q0 = cc[0] + a, q1 = cc[1] + b, q2 = cc[2] + c;
          if (g.periodic[0]) q0 = (q0 + g.n[0]) % g.n[0];
          else if (q0 < 0 || q0 >= g.n[0]) continue;
          if (g.periodic[1]) q1 = (q1 + g.n[1]) % g.n[1];
          else if (q1 < 0 || q1 >= g.n[1]) continue;
          if (g.periodic[2]) q2 = (q2 + g.n[2]) % g.n[2];
          else if (q2 < 0 || q2 >= g.n[2]) continue;
          int cid = (q0 * g.n[1] + q1) * g.n[2] + q2;
          int s = start[cid], e = start[cid + 1];
          for (int base = s; base < e; base += 64) {
            int p = base + lane;
            bool ok = false;
            int j = -1;
            if (p < e) {
              j = cell_atoms[p];
              if (j != i && handled[j]) {
                double dx = mic(xi - x[j], bx.L[0]);
                double dy = mic(yi - y[j], bx.L[1]);
                double dz = mic(zi - z[j], bx.L[2]);
                double d2 = dx * dx + dy * dy + dz * dz;
                ok = d2 < range2;
              }
            }
            unsigned long long mask = __ballot(ok);
            int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (ok) {
              int slot = m + before;
              if (slot < CAP) my[slot] = j;
            }
            m += __popcll(mask);
          }
        }
  }
  if (m > CAP) {
    if (lane == 0) raise(err, CAP >= 256 ? ERR_LONG_NL_OVERFLOW : ERR_SHORT_NL_OVERFLOW);
    m = CAP;
  }
  // pad to a power of two and sort ascending
  int p2 = 1;
  while (p2 < m) p2 <<= 1;
  for (int q = m + lane; q < p2; q += 64) my[q] = 0x7fffffff;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int k = 2; k <= p2; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int q = lane; q < p2; q += 64) {
        int partner = q ^ jj;
        if (partner > q) {
          int a = my[q], b = my[partner];
          bool up = ((q & k) == 0);
          if ((a > b) == up) {
            my[q] = b;
            my[partner] = a;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  int stride = CAP;
  for (int q = lane; q < m; q += 64) list[(size_t)i * stride + q] = my[q];
  if (lane == 0) cnt[i] = m;
}

template __global__ void k_nl_build<64>(const int*, int, const double*, const double*, const double*,
                                        const unsigned char*, Box, CellGrid, double, const int*, const int*,
                                        const int*, int*, int*, double*, unsigned*);
template __global__ void k_nl_build<512>(const int*, int, const double*, const double*, const double*,
                                         const unsigned char*, Box, CellGrid, double, const int*, const int*,
                                         const int*, int*, int*, double*, unsigned*);

// Top-2 accumulated displacement per neighbour list, NeighbourList.cxx:108-133.
// Stage 1: per-block top-2; stage 2: one block merges and sets the flags.
__device__ __forceinline__ void top2_insert(double v, double& d1, double& d2) {
  if (v >= d1) {
    d2 = d1;
    d1 = v;
  } else if (v >= d2) {
    d2 = v;
  }
}

__global__ void k_disp_top2(int n, const double* __restrict__ ds, const double* __restrict__ dl,
                            double* __restrict__ part) {
  __shared__ double s1[256], s2[256], l1[256], l2[256];
  double a1 = 0, a2 = 0, b1 = 0, b2 = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double m = sqrt(ds[3 * i] * ds[3 * i] + ds[3 * i + 1] * ds[3 * i + 1] + ds[3 * i + 2] * ds[3 * i + 2]);
    top2_insert(m, a1, a2);
    double q = sqrt(dl[3 * i] * dl[3 * i] + dl[3 * i + 1] * dl[3 * i + 1] + dl[3 * i + 2] * dl[3 * i + 2]);
    top2_insert(q, b1, b2);
  }
  int t = threadIdx.x;
  s1[t] = a1; s2[t] = a2; l1[t] = b1; l2[t] = b2;
  __syncthreads();
  if (t == 0) {
    double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
    for (int q = 0; q < (int)blockDim.x; q++) {
      top2_insert(s1[q], x1, x2);
      top2_insert(s2[q], x1, x2);
      top2_insert(l1[q], y1, y2);
      top2_insert(l2[q], y1, y2);
    }
    part[4 * blockIdx.x + 0] = x1;
    part[4 * blockIdx.x + 1] = x2;
    part[4 * blockIdx.x + 2] = y1;
    part[4 * blockIdx.x + 3] = y2;
  }
}

__global__ void k_disp_decide(int nblk, const double* __restrict__ part, double lim_s, double lim_l,
                              int* __restrict__ flags, long long* __restrict__ rebuilds) {
  if (threadIdx.x != 0) return;
  double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
  for (int q = 0; q < nblk; q++) {
    top2_insert(part[4 * q], x1, x2);
    top2_insert(part[4 * q + 1], x1, x2);
    top2_insert(part[4 * q + 2], y1, y2);
    top2_insert(part[4 * q + 3], y1, y2);
  }
  flags[0] = (x1 + x2 > lim_s) ? 1 : 0;
  flags[1] = (y1 + y2 > lim_l) ? 1 : 0;
  if (flags[0]) rebuilds[0] += 1;
  if (flags[1]) rebuilds[1] += 1;
}

__global__ void k_cell_clear(const int* __restrict__ flag, int ncell, int* __restrict__ cnt) {
  if (!*flag) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += gridDim.x * blockDim.x) cnt[i] = 0;
}

// ----------------------------------------------------- coordination pass
// REBO::countNeighbours (REBO.cxx:328-393) plus the per-atom bond list: all
// short-list neighbours with r <= R2 in list order, with r, f, df and the
// minimum-image vector.
__global__ void k_coord(Atoms A, const Tables* __restrict__ T, Box bx, unsigned* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  int ti = A.type[i];
  double Nt = 0, NH = 0, NC = 0;
  int nb = 0;
  if (ti <= T_H) {
    int cnt = A.nls_cnt[i];
    const int* L = A.nls + (size_t)i * A.cap_s;
    for (int q = 0; q < cnt; q++) {
      int k = L[q];
      int tk = A.type[k];
      D3 rv = depos(A.x, A.y, A.z, i, k, bx);
      double r = sqrt(rv.x * rv.x + rv.y * rv.y + rv.z * rv.z);
      double R2 = T->rebo.R[ti][tk][1];
      if (r > R2) continue;
      double R1 = T->rebo.R[ti][tk][0];
      double f, df;
      cutoff(r, R1, R2, f, df);
      Nt += f;
      if (tk == T_H) NH += f;
      if (tk == T_C) NC += f;
      if (nb < A.maxb) {
        int o = i * A.maxb + nb;
        A.bj[o] = k;
        A.br[o] = r;
        A.bf[o] = f;
        A.bdf[o] = df;
        A.bx[o] = rv.x;
        A.by[o] = rv.y;
        A.bz[o] = rv.z;
      } else {
        raise(err, ERR_BOND_OVERFLOW);
      }
      nb++;
    }
  }
  A.Nt[i] = Nt;
  A.NH[i] = NH;
  A.NC[i] = NC;
  A.bcnt[i] = nb < A.maxb ? nb : A.maxb;
}

// Bonds (i<j) in the order of i's list, appended to one array.
__global__ void k_bond_enum(Atoms A, int2* __restrict__ out, int* __restrict__ count, int cap, unsigned* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  int nb = A.bcnt[i];
  int mine = 0;
  for (int s = 0; s < nb; s++) mine += (A.bj[i * A.maxb + s] > i);
  if (!mine) return;
  int base = atomicAdd(count, mine);
  for (int s = 0; s < nb; s++) {
    int j = A.bj[i * A.maxb + s];
    if (j > i) {
      if (base < cap) out[base] = make_int2(i, j);
      else raise(err, ERR_QUEUE_OVERFLOW);
      base++;
    }
  }
}

// ------------------------------------------------------- bond order b_ij
// The pair (i,j) whose bond order is needed: a REBO bond, or an LJ pair for
// b* (then r_eff = R0 inside ExpTerm only, AIREBO_LJ.cxx:191-202).
struct BPair {
  int i, j, ti, tj;
  D3 rv;      // x_i - x_j
  double r;
  double f;   // REBO cutoff value of (i,j) (0 for b* pairs beyond R2)
  double reff;
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

// Geometry of the k-th neighbour of side s (s=0: centre i, partner j; s=1:
// centre j, partner i).
struct Side {
  int a, b, ta, tb;
  D3 rv;  // x_a - x_b
};
__device__ __forceinline__ Side side_of(const BPair& p, int s) {
  Side q;
  if (s == 0) {
    q.a = p.i; q.b = p.j; q.ta = p.ti; q.tb = p.tj; q.rv = p.rv;
  } else {
    q.a = p.j; q.b = p.i; q.ta = p.tj; q.tb = p.ti; q.rv = -p.rv;
  }
  return q;
}

// G(cos theta_bak) with its partial derivatives (REBO.cxx:277-326).
struct GVal {
  double G, dGdc, dGdN;
  bool has_dc, has_dN;
};
__device__ __forceinline__ GVal gfun(const ReboParams& P, int ta, double c, double N) {
  GVal g;
  if (ta == T_C) {
    double Sp = sprime(N, P.Ntot[0], P.Ntot[1]);
    double g1, g1d, g2, g2d;
    gspline(P.gC1, c, g1, g1d);
    gspline(P.gC2, c, g2, g2d);
    g.G = g2 + Sp * (g1 - g2);
    g.has_dc = (g2d != 0.0 || g1d != 0.0);
    g.dGdc = g2d + g1d * Sp + (-g2d * Sp);
    double dSp = dsprime(N, P.Ntot[0], P.Ntot[1]);
    g.has_dN = ((g2 != 0.0 || g1 != 0.0) && dSp != 0.0);
    g.dGdN = g1 * dSp + (-g2 * dSp);
  } else {
    double v, d;
    gspline(P.gH, c, v, d);
    g.G = v;
    g.has_dc = (d != 0.0);
    g.dGdc = d;
    g.has_dN = false;
    g.dGdN = 0.0;
  }
  return g;
}

__device__ __forceinline__ double exp_term(const ReboParams& P, const Side& q, int tk, double r_ak, double reff_ab) {
  if (q.ta != T_H) return 1.0;
  return exp((4.0 / kAo) * ((P.rho[q.ta][tk] - r_ak) - (P.rho[q.ta][q.tb] - reff_ab)));
}

__device__ __forceinline__ double cos_theta(D3 a, D3 b, double ra, double rb, unsigned* err, double& raw) {
  double e1 = dot(a, b);
  double e2 = ra * rb;
  double c = e1 / e2;
  raw = c;
  if (!(c >= -1.0 - 1e-7 && c <= 1.0 + 1e-7)) raise(err, ERR_COS_RANGE);
  if (c < -1.0) c = -1.0;
  if (c > 1.0) c = 1.0;
  return c;
}

// Forward pass of b_ij = (p_ij + p_ji)/2 + pi_rc + pi_dh (REBO::Baver,
// REBO.cxx:200-211, with D :235-262, p_sigma_pi :264-275, Nconj :540-624,
// P :626-657, pi_rc :659-714, pi_dh :716-814).
__device__ void bo_forward(const Atoms& A, const Tables* __restrict__ T, const BPair& p, BOState& st,
                           unsigned* err) {
  const ReboParams& P = T->rebo;
  double R2ij = P.R[p.ti][p.tj][1];
  // coordination numbers excluding the partner
  st.N[0] = coord_excl(A.Nt[p.i], true, p.r, R2ij, p.f, err);
  st.N[1] = coord_excl(A.Nt[p.j], true, p.r, R2ij, p.f, err);
  st.NHc[0] = coord_excl(A.NH[p.i], p.tj == T_H, p.r, R2ij, p.f, err);
  st.NCc[0] = coord_excl(A.NC[p.i], p.tj == T_C, p.r, R2ij, p.f, err);
  st.NHc[1] = coord_excl(A.NH[p.j], p.ti == T_H, p.r, R2ij, p.f, err);
  st.NCc[1] = coord_excl(A.NC[p.j], p.ti == T_C, p.r, R2ij, p.f, err);
  for (int s = 0; s < 2; s++) {
    Side q = side_of(p, s);
    double D = 1.0;
    int nb = A.bcnt[q.a];
    for (int t = 0; t < nb; t++) {
      BondRef k = bond_at(A, q.a, t);
      if (k.k == q.b) continue;
      int tk = A.type[k.k];
      double E = exp_term(P, q, tk, k.r, p.reff);
      if (E != 0.0) {
        double raw;
        double c = cos_theta(q.rv, k.rv, p.r, k.r, err, raw);
        GVal g = gfun(P, q.ta, c, st.N[s]);
        D += g.G * k.f * E;
      }
    }
    st.P[s][0] = st.P[s][1] = st.P[s][2] = 0.0;
    if (q.ta == T_C && q.tb <= T_H) bicubic(T->P[q.tb == T_C ? 0 : 1], st.NHc[s], st.NCc[s], st.P[s]);
    D += st.P[s][0];
    st.D[s] = D;
  }
  double pij = pow(st.D[0], -0.5);
  double pji = pow(st.D[1], -0.5);
  // Nconj partial sums: carbon neighbours k of the centre, f_ka * S'(N_ka)
  for (int s = 0; s < 2; s++) {
    Side q = side_of(p, s);
    double sum = 0.0;
    int nb = A.bcnt[q.a];
    for (int t = 0; t < nb; t++) {
      BondRef k = bond_at(A, q.a, t);
      if (k.k == q.b) continue;
      int tk = A.type[k.k];
      if (tk != T_C) continue;
      double Nk = coord_excl(A.Nt[k.k], true, k.r, P.R[tk][q.ta][1], k.f, err);
      sum += k.f * sprime(Nk, P.Ntot[0], P.Ntot[1]);
    }
    st.sum[s] = sum;
  }
  st.Nconj = 1.0 + st.sum[0] * st.sum[0] + st.sum[1] * st.sum[1];
  // pi_rc by pair type; the H-C case evaluates F_CH(N_j, N_i, Nconj_ji)
  for (int q = 0; q < 4; q++) st.rc[q] = 0.0;
  if (p.ti == T_C && p.tj == T_C) {
    tricubic<16>(T->T3[0], st.N[0], st.N[1], st.Nconj, 15, 15, st.rc);
  } else if (p.ti == T_C && p.tj == T_H) {
    tricubic<16>(T->T3[1], st.N[0], st.N[1], st.Nconj, 15, 15, st.rc);
  } else if (p.ti == T_H && p.tj == T_C) {
    double nconj_ji = 1.0 + st.sum[1] * st.sum[1] + st.sum[0] * st.sum[0];
    double o[4];
    tricubic<16>(T->T3[1], st.N[1], st.N[0], nconj_ji, 15, 15, o);
    st.rc[0] = o[0];
    st.rc[1] = o[2];  // d/dN_i
    st.rc[2] = o[1];  // d/dN_j
    st.rc[3] = o[3];
  } else if (p.ti == T_H && p.tj == T_H) {
    tricubic<16>(T->T3[2], st.N[0], st.N[1], st.Nconj, 15, 15, st.rc);
  }
  // pi_dh: C-C only, skipped when T_ij == 0 (REBO.cxx:802-814)
  st.T[0] = st.T[1] = st.T[2] = st.T[3] = 0.0;
  st.dh = 0.0;
  double pidh = 0.0;
  if (p.ti == T_C && p.tj == T_C) {
    tricubic<16>(T->T3[3], st.N[0], st.N[1], st.Nconj, 15, 15, st.T);
    if (st.T[0] != 0.0) {
      double s = 0.0;
      int nbi = A.bcnt[p.i], nbj = A.bcnt[p.j];
      for (int a = 0; a < nbi; a++) {
        BondRef k = bond_at(A, p.i, a);
        if (k.k == p.j) continue;
        for (int b = 0; b < nbj; b++) {
          BondRef l = bond_at(A, p.j, b);
          if (l.k == p.i) continue;
          if (k.k == l.k) continue;
          double sin_ik = sqrt(norm2(cross(p.rv, k.rv))) / (p.r * k.r);
          if (fabs(sin_ik) < 0.1) continue;
          int tk = A.type[k.k], tl = A.type[l.k];
          double fik = sprime(k.r, P.Rp[tk][p.ti][0], P.Rp[tk][p.ti][1]);
          double sin_jl = sqrt(norm2(cross(p.rv, l.rv))) / (p.r * l.r);
          if (fabs(sin_jl) < 0.1) continue;
          double fjl = sprime(l.r, P.Rp[tl][p.tj][0], P.Rp[tl][p.tj][1]);
          D3 Pp = cross(-p.rv, k.rv), Q = cross(p.rv, l.rv);
          double e2 = sqrt(norm2(Pp)) * sqrt(norm2(Q));
          double cd = (e2 == 0.0) ? 0.0 : dot(Pp, Q) / e2;
          if (cd < -1.0) cd = -1.0;
          if (cd > 1.0) cd = 1.0;
          double c = -cd;
          s += (1.0 - c * c) * fik * fjl;
        }
      }
      st.dh = s;
      pidh = st.T[0] * s;
    }
  }
  st.b = (pij + pji) / 2.0 + st.rc[0] + pidh;
}

// Scatter c * d f_ab / dx for the pair (a,k) along u_ak = rv/r (a +=, k -=).
__device__ __forceinline__ void pair_push(double* g, int a, int k, D3 rv, double r, double c, D3& ga) {
  D3 v = rv * (c / r);
  ga = ga + v;
  atomic_add3(g, k, -v);
}

// Adjoint pass: adds W * db_ij/dx to grad.
__device__ void bo_backward(const Atoms& A, const Tables* __restrict__ T, const BPair& p, const BOState& st,
                            double W, double* __restrict__ grad, unsigned* err) {
  const ReboParams& P = T->rebo;
  D3 gi = d3(0, 0, 0), gj = d3(0, 0, 0);
  double dN[2] = {0, 0}, dNH[2] = {0, 0}, dNC[2] = {0, 0}, dNconj = 0.0;
  // sigma-pi terms
  for (int s = 0; s < 2; s++) {
    Side q = side_of(p, s);
    D3& ga = (s == 0) ? gi : gj;
    D3& gb = (s == 0) ? gj : gi;
    double seedD = (-0.5 * pow(st.D[s], -0.5 - 1.0)) * (W / 2.0);
    int nb = A.bcnt[q.a];
    for (int t = 0; t < nb; t++) {
      BondRef k = bond_at(A, q.a, t);
      if (k.k == q.b) continue;
      int tk = A.type[k.k];
      double E = exp_term(P, q, tk, k.r, p.reff);
      if (E == 0.0) continue;
      double raw;
      double c = cos_theta(q.rv, k.rv, p.r, k.r, err, raw);
      GVal g = gfun(P, q.ta, c, st.N[s]);
      double V1 = k.f * E * seedD;
      D3 gk = d3(0, 0, 0);
      if (g.has_dc) {
        double cc = g.dGdc * V1;
        double e2 = p.r * k.r;
        D3 da = k.rv / e2 - q.rv * (raw / (p.r * p.r));
        D3 db = q.rv / e2 - k.rv * (raw / (k.r * k.r));
        ga = ga + (da + db) * cc;
        gb = gb - da * cc;
        gk = gk - db * cc;
      }
      if (g.has_dN) dN[s] += g.dGdN * V1;
      // d f_ak
      if (k.df != 0.0) {
        D3 v = k.rv * (k.df * (g.G * E * seedD) / k.r);
        ga = ga + v;
        gk = gk - v;
      }
      if (q.ta == T_H) {
        double dt = E * (4.0 / kAo);
        double V3 = g.G * k.f * seedD;
        // r_ab (or the b* pair's r_alter, whose derivative is the unit vector)
        D3 u = q.rv * (dt * V3 / p.r);
        ga = ga + u;
        gb = gb - u;
        D3 v = k.rv * (-dt * V3 / k.r);
        ga = ga + v;
        gk = gk - v;
      }
      atomic_add3(grad, k.k, gk);
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
    double Ws = st.T[0] * W;
    int nbi = A.bcnt[p.i], nbj = A.bcnt[p.j];
    for (int a = 0; a < nbi; a++) {
      BondRef k = bond_at(A, p.i, a);
      if (k.k == p.j) continue;
      for (int b = 0; b < nbj; b++) {
        BondRef l = bond_at(A, p.j, b);
        if (l.k == p.i) continue;
        if (k.k == l.k) continue;
        double sin_ik = sqrt(norm2(cross(p.rv, k.rv))) / (p.r * k.r);
        if (fabs(sin_ik) < 0.1) continue;
        int tk = A.type[k.k], tl = A.type[l.k];
        double R1k = P.Rp[tk][p.ti][0], R2k = P.Rp[tk][p.ti][1];
        double fik = sprime(k.r, R1k, R2k);
        double sin_jl = sqrt(norm2(cross(p.rv, l.rv))) / (p.r * l.r);
        if (fabs(sin_jl) < 0.1) continue;
        double R1l = P.Rp[tl][p.tj][0], R2l = P.Rp[tl][p.tj][1];
        double fjl = sprime(l.r, R1l, R2l);
        D3 a_ = p.rv, b_ = k.rv, c_ = l.rv;
        D3 Pp = cross(-a_, b_), Q = cross(a_, c_);
        double p2 = norm2(Pp), q2 = norm2(Q);
        double e2 = sqrt(p2) * sqrt(q2);
        double cs = (e2 == 0.0) ? 0.0 : dot(Pp, Q) / e2;
        double cd = cs;
        if (cd < -1.0) cd = -1.0;
        if (cd > 1.0) cd = 1.0;
        double c = -cd;
        D3 gk = d3(0, 0, 0), gl = d3(0, 0, 0);
        double V = 2.0 * c * fik * fjl * Ws;
        if (V != 0.0 && e2 != 0.0) {
          D3 gP = Q / e2 - Pp * (cs / p2);
          D3 gQ = Pp / e2 - Q * (cs / q2);
          D3 da = cross(gP, b_) + cross(c_, gQ);
          D3 db = cross(a_, gP);
          D3 dc = cross(gQ, a_);
          gi = gi + (da + db) * V;
          gj = gj + (dc - da) * V;
          gk = gk - db * V;
          gl = gl - dc * V;
        }
        double Vk = (1.0 - c * c) * fjl * Ws;
        if (Vk != 0.0 && k.r >= R1k && k.r <= R2k) {
          D3 v = k.rv * (dsprime(k.r, R1k, R2k) * Vk / k.r);
          gi = gi + v;
          gk = gk - v;
        }
        double Vl = (1.0 - c * c) * fik * Ws;
        if (Vl != 0.0 && l.r >= R1l && l.r <= R2l) {
          D3 v = l.rv * (dsprime(l.r, R1l, R2l) * Vl / l.r);
          gj = gj + v;
          gl = gl - v;
        }
        atomic_add3(grad, k.k, gk);
        atomic_add3(grad, l.k, gl);
      }
    }
  }
  // coordination derivatives (dfThem, REBO.cxx:408-421) and Nconj
  for (int s = 0; s < 2; s++) {
    Side q = side_of(p, s);
    D3& ga = (s == 0) ? gi : gj;
    double sconj = (st.sum[s] != 0.0) ? 2.0 * st.sum[s] * dNconj : 0.0;
    int nb = A.bcnt[q.a];
    for (int t = 0; t < nb; t++) {
      BondRef k = bond_at(A, q.a, t);
      if (k.k == q.b) continue;
      int tk = A.type[k.k];
      double c = dN[s];
      if (tk == T_H) c += dNH[s];
      if (tk == T_C) c += dNC[s];
      double Sp = 0.0, coefk = 0.0;
      if (tk == T_C && sconj != 0.0) {
        double Nk = coord_excl(A.Nt[k.k], true, k.r, P.R[tk][q.ta][1], k.f, err);
        Sp = sprime(Nk, P.Ntot[0], P.Ntot[1]);
        c += Sp * sconj;
        coefk = k.f * dsprime(Nk, P.Ntot[0], P.Ntot[1]) * sconj;
      }
      if (c != 0.0 && k.df != 0.0) pair_push(grad, q.a, k.k, k.rv, k.r, k.df * c, ga);
      if (coefk != 0.0) {
        // d N_ka: neighbours m of k other than a
        D3 gk = d3(0, 0, 0);
        int nbk = A.bcnt[k.k];
        for (int u = 0; u < nbk; u++) {
          BondRef m = bond_at(A, k.k, u);
          if (m.k == q.a || m.df == 0.0) continue;
          D3 v = m.rv * (m.df * coefk / m.r);
          gk = gk + v;
          atomic_add3(grad, m.k, -v);
        }
        atomic_add3(grad, k.k, gk);
      }
    }
  }
  atomic_add3(grad, p.i, gi);
  atomic_add3(grad, p.j, gj);
}

// ------------------------------------------------------------ REBO bonds
// REBO::operator() (REBO.cxx:32-65): E += VR + b_ij * VA for every bond.
__global__ void __launch_bounds__(256) k_rebo_bonds(Atoms A, const Tables* __restrict__ T, Box bx,
                                                   const int2* __restrict__ bonds, const int* __restrict__ nbonds,
                                                   double* __restrict__ grad, double* __restrict__ eacc,
                                                   unsigned* err) {
  double e = 0.0;
  int nbt = min(*nbonds, A.bond_cap);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nbt; t += gridDim.x * blockDim.x) {
    const ReboParams& P = T->rebo;
    int2 bd = bonds[t];
    BPair p;
    p.i = bd.x;
    p.j = bd.y;
    p.ti = A.type[p.i];
    p.tj = A.type[p.j];
    p.rv = depos(A.x, A.y, A.z, p.i, p.j, bx);
    p.r = sqrt(norm2(p.rv));
    double f, df;
    cutoff(p.r, P.R[p.ti][p.tj][0], P.R[p.ti][p.tj][1], f, df);
    p.f = f;
    p.reff = p.r;
    int a = p.ti, b = p.tj;
    double r = p.r;
    // VR (REBO.cxx:132-160) and VA (:162-198)
    double VR = 0.0, VA = 0.0, cr = 0.0;
    if (f != 0.0) {
      double t1 = P.A[a][b] * exp(-P.alpha[a][b] * r);
      double val = (1 + P.Q[a][b] / r) * t1;
      double der = -t1 * (P.Q[a][b] + P.alpha[a][b] * r * r + P.alpha[a][b] * r * P.Q[a][b]) / (r * r);
      VR = f * val;
      cr += der * f + val * df;
      double u1 = P.B[0][a][b] * exp(-P.beta[0][a][b] * r);
      double u2 = P.B[1][a][b] * exp(-P.beta[1][a][b] * r);
      double u3 = P.B[2][a][b] * exp(-P.beta[2][a][b] * r);
      double vsum = u1 + u2 + u3;
      VA = -f * vsum;
      e += VR;
      if (VA != 0.0) {
        BOState st;
        bo_forward(A, T, p, st, err);
        double bb = st.b;
        bo_backward(A, T, p, st, VA, grad, err);
        double vder = (-P.beta[0][a][b]) * u1 + (-P.beta[1][a][b]) * u2 + (-P.beta[2][a][b]) * u3;
        cr += (-vder * f * bb) + (-vsum * bb) * df;
        e += bb * VA;
      }
      D3 v = p.rv * (cr / r);
      atomic_add3(grad, p.i, v);
      atomic_add3(grad, p.j, -v);
    }
  }
  block_accumulate(e, eacc);
}

// ------------------------------------------------------------ C_ij paths
// Best path weight over <=3 REBO bonds from each atom (AIREBO::fill_Cij,
// AIREBO_LJ.cxx:289-358), enumerated in the reference's loop order with its
// tie rules: a later 1-hop replaces an equal-weight longer path, a later
// 2-hop replaces an equal 3-hop, a 3-hop needs a strictly larger weight.
// Stored per atom sorted by target for the LJ merge.
__global__ void k_cij_build(Atoms A, unsigned* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  int cap = A.cap_c;
  int* tg = A.c_tgt + (size_t)i * cap;
  double* w = A.c_w + (size_t)i * cap;
  int* km = A.c_km + (size_t)i * cap;  // packed (hops, k, m): hops in bits 30-31
  int cnt = 0;
  auto update = [&](int target, double nw, int hops, int k, int m) {
    if (target == i) return;
    int e = -1;
    for (int q = 0; q < cnt; q++)
      if (tg[q] == target) {
        e = q;
        break;
      }
    if (e < 0) {
      if (!(nw > 0.0)) return;
      if (cnt >= cap) {
        raise(err, ERR_QUEUE_OVERFLOW);
        return;
      }
      e = cnt++;
      tg[e] = target;
      w[e] = nw;
      A.c_k[(size_t)i * cap + e] = k;
      km[e] = (hops << 28) | (m & 0x0fffffff);
      return;
    }
    double ow = w[e];
    int oh = km[e] >> 28;
    if (nw > ow || (ow > 0.0 && nw == ow && oh > hops && hops < 3)) {
      w[e] = nw;
      A.c_k[(size_t)i * cap + e] = k;
      km[e] = (hops << 28) | (m & 0x0fffffff);
    }
  };
  if (A.type[i] <= T_H) {
    int nb1 = A.bcnt[i];
    for (int s1 = 0; s1 < nb1; s1++) {
      int k = A.bj[i * A.maxb + s1];
      double f1 = A.bf[i * A.maxb + s1];
      update(k, f1, 1, -1, 0);
      int nb2 = A.bcnt[k];
      for (int s2 = 0; s2 < nb2; s2++) {
        int m = A.bj[k * A.maxb + s2];
        double f2 = A.bf[k * A.maxb + s2];
        update(m, f1 * f2, 2, k, 0);
        int nb3 = A.bcnt[m];
        for (int s3 = 0; s3 < nb3; s3++) {
          int q = A.bj[m * A.maxb + s3];
          double f3 = A.bf[m * A.maxb + s3];
          update(q, f1 * f2 * f3, 3, k, m);
        }
      }
    }
  }
  // insertion sort by target
  for (int a = 1; a < cnt; a++) {
    int t0 = tg[a];
    double w0 = w[a];
    int k0 = A.c_k[(size_t)i * cap + a];
    int km0 = km[a];
    int b = a - 1;
    while (b >= 0 && tg[b] > t0) {
      tg[b + 1] = tg[b];
      w[b + 1] = w[b];
      A.c_k[(size_t)i * cap + b + 1] = A.c_k[(size_t)i * cap + b];
      km[b + 1] = km[b];
      b--;
    }
    tg[b + 1] = t0;
    w[b + 1] = w0;
    A.c_k[(size_t)i * cap + b + 1] = k0;
    km[b + 1] = km0;
  }
  A.c_cnt[i] = cnt;
}

// Gradient of C_ij = 1 - prod f over the stored path, seeded with V
// (AIREBO::Cij, AIREBO_LJ.cxx:382-408).
__device__ void cij_backward(const Atoms& A, const Tables* __restrict__ T, Box bx, int i, int j, int e, double V,
                             double* __restrict__ grad) {
  size_t o = (size_t)i * A.cap_c + e;
  int hops = A.c_km[o] >> 28;
  int k = A.c_k[o];
  int m = A.c_km[o] & 0x0fffffff;
  int nodes[4];
  int nn = 0;
  nodes[nn++] = i;
  if (hops >= 2) nodes[nn++] = k;
  if (hops >= 3) nodes[nn++] = m;
  nodes[nn++] = j;
  double f[3], df[3], r[3];
  D3 rv[3];
  for (int q = 0; q < hops; q++) {
    int a = nodes[q], b = nodes[q + 1];
    rv[q] = depos(A.x, A.y, A.z, a, b, bx);
    r[q] = sqrt(norm2(rv[q]));
    int ta = A.type[a], tb = A.type[b];
    cutoff(r[q], T->rebo.R[ta][tb][0], T->rebo.R[ta][tb][1], f[q], df[q]);
  }
  for (int q = 0; q < hops; q++) {
    if (df[q] == 0.0) continue;
    double mm = 1.0;
    for (int u = 0; u < hops; u++)
      if (u != q) mm *= f[u];
    double c = df[q] * (-mm * V);
    D3 v = rv[q] * (c / r[q]);
    atomic_add3(grad, nodes[q], v);
    atomic_add3(grad, nodes[q + 1], -v);
  }
}

// ---------------------------------------------------------------- LJ
__device__ __forceinline__ double Sfun(double a, double mn, double mx) {
  if (a < mn) return 1.0;
  if (a > mx) return 0.0;
  double t = (a - mn) / (mx - mn);
  return 1.0 - t * t * (3.0 - 2.0 * t);
}
__device__ __forceinline__ double dSfun(double a, double mn, double mx) {
  if (a < mn || a > mx) return 0.0;
  double t = (a - mn) / (mx - mn);
  return (-2.0 * t * (3.0 - 2.0 * t) + 2.0 * t * t) / (mx - mn);
}

// AIREBO::operator() (AIREBO_LJ.cxx:34-77) for the pairs that do not need a
// b* bond order; the others are queued for k_bstar.  One thread per atom i
// walks its sorted long list (j > i) and its sorted C_ij table together.
__global__ void __launch_bounds__(256) k_lj_pairs(Atoms A, const Tables* __restrict__ T, Box bx,
                                                 double* __restrict__ grad, double* __restrict__ eacc,
                                                 int4* __restrict__ queue, int* __restrict__ qcount, int qcap,
                                                 long long* __restrict__ npairs, unsigned* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  long long np = 0;
  if (i < A.n && A.type[i] <= T_H) {
    const LJParams& L = T->lj;
    int ti = A.type[i];
    int cnt = A.nll_cnt[i];
    const int* list = A.nll + (size_t)i * A.cap_l;
    int ccnt = A.c_cnt[i];
    const int* ctg = A.c_tgt + (size_t)i * A.cap_c;
    int cp = 0;
    D3 gi = d3(0, 0, 0);
    double xi = A.x[i], yi = A.y[i], zi = A.z[i];
    for (int q = 0; q < cnt; q++) {
      int j = list[q];
      if (j <= i) continue;
      int tj = A.type[j];
      if (tj > T_H) continue;
      D3 rv = d3(mic(xi - A.x[j], bx.L[0]), mic(yi - A.y[j], bx.L[1]), mic(zi - A.z[j], bx.L[2]));
      double d2 = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
      if (d2 > L.R2sq[ti][tj]) continue;
      np++;
      while (cp < ccnt && ctg[cp] < j) cp++;
      int ce = (cp < ccnt && ctg[cp] == j) ? cp : -1;
      double w = (ce >= 0) ? A.c_w[(size_t)i * A.cap_c + ce] : 0.0;
      double C = 1 - w;
      if (!(C > 0.0)) continue;
      double r = sqrt(d2);
      double Str = Sfun(r, L.RLJ[ti][tj][0], L.RLJ[ti][tj][1]);
      if (Str != 0.0) {
        int slot = atomicAdd(qcount, 1);
        if (slot < qcap) queue[slot] = make_int4(i, j, ce, 0);
        else raise(err, ERR_QUEUE_OVERFLOW);
        continue;
      }
      double f, df;
      cutoff(r, L.R[ti][tj][0], L.R[ti][tj][1], f, df);
      if (f == 0.0) continue;
      double s = L.sigma[ti][tj] / r;
      double s6 = s;
      for (int u = 2; u <= 6; u++) s6 *= s;
      double s12 = s6 * s6;
      double z = L.zeta[ti][tj];
      double val = 4.0 * z * (s12 - s6);
      double der = 4.0 * z * (-12.0 * s12 / r + 6.0 * s6 / r);
      double LJ = f * val;
      // VLJ seeded with SrSb*C (SrSb = 1 here), Cij seeded with LJ
      double V = 1.0 * C;
      double cr = der * f * V + val * V * df;
      D3 v = rv * (cr / r);
      gi = gi + v;
      atomic_add3(grad, j, -v);
      if (ce >= 0) cij_backward(A, T, bx, i, j, ce, 1.0 * LJ, grad);
      e += 1.0 * C * LJ;
    }
    atomic_add3(grad, i, gi);
  }
  block_accumulate(e, eacc);
  if (np) atomicAdd((unsigned long long*)npairs, (unsigned long long)np);
}

// LJ pairs that need b*_ij (S_r != 0): the whole StrStb / VLJ / Cij chain.
__global__ void __launch_bounds__(256) k_bstar(Atoms A, const Tables* __restrict__ T, Box bx,
                                              const int4* __restrict__ queue, const int* __restrict__ qcount,
                                              int qcap, double* __restrict__ grad, double* __restrict__ eacc,
                                              unsigned* err) {
  double e = 0.0;
  int nq = min(*qcount, qcap);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nq; t += gridDim.x * blockDim.x) {
    const LJParams& L = T->lj;
    const ReboParams& P = T->rebo;
    int4 qd = queue[t];
    int i = qd.x, j = qd.y, ce = qd.z;
    int ti = A.type[i], tj = A.type[j];
    D3 rv = depos(A.x, A.y, A.z, i, j, bx);
    double r = sqrt(norm2(rv));
    double w = (ce >= 0) ? A.c_w[(size_t)i * A.cap_c + ce] : 0.0;
    double C = 1 - w;
    double rmin = L.RLJ[ti][tj][0], rmax = L.RLJ[ti][tj][1];
    double Str = Sfun(r, rmin, rmax);
    // b* pair: REBO cutoffs at the actual geometry, r_alter = R0 in ExpTerm
    BPair p;
    p.i = i;
    p.j = j;
    p.ti = ti;
    p.tj = tj;
    p.rv = rv;
    p.r = r;
    double fb, dfb;
    cutoff(r, P.R[ti][tj][0], P.R[ti][tj][1], fb, dfb);
    p.f = fb;
    p.reff = P.R[ti][tj][0];
    BOState st;
    bo_forward(A, T, p, st, err);
    double bmin = L.b[ti][tj][0], bmax = L.b[ti][tj][1];
    double Stb = Sfun(st.b, bmin, bmax);
    double SrSb = 1.0 + (Str * Stb - Str);
    double f, df;
    cutoff(r, L.R[ti][tj][0], L.R[ti][tj][1], f, df);
    double s = L.sigma[ti][tj] / r;
    double s6 = s;
    for (int u = 2; u <= 6; u++) s6 *= s;
    double s12 = s6 * s6;
    double z = L.zeta[ti][tj];
    double val = 4.0 * z * (s12 - s6);
    double der = 4.0 * z * (-12.0 * s12 / r + 6.0 * s6 / r);
    double LJ = (f == 0.0) ? 0.0 : f * val;
    double cr = 0.0;
    if (f != 0.0) {
      double V = SrSb * C;
      cr += der * f * V + val * V * df;
    }
    // StrStb seeded with C*LJ
    double V2 = C * LJ;
    if (V2 != 0.0) {
      double dStr = dSfun(r, rmin, rmax);
      if (dStr != 0.0) cr += -dStr * V2;
      double dStb = dSfun(st.b, bmin, bmax);
      if (dStr != 0.0 && Stb != 0.0) cr += dStr * Stb * V2;
      if (dStb != 0.0) bo_backward(A, T, p, st, Str * dStb * V2, grad, err);
    }
    D3 v = rv * (cr / r);
    atomic_add3(grad, i, v);
    atomic_add3(grad, j, -v);
    if (ce >= 0) cij_backward(A, T, bx, i, j, ce, SrSb * LJ, grad);
    e += SrSb * C * LJ;
  }
  block_accumulate(e, eacc);
}

// ------------------------------------------------------------ torsion
// ETors::operator() (AIREBO_ETors.cxx:32-115) over C-C bonds.
__global__ void __launch_bounds__(256) k_etors(Atoms A, const Tables* __restrict__ T, Box bx,
                                              const int2* __restrict__ bonds, const int* __restrict__ nbonds,
                                              int compat, double* __restrict__ grad, double* __restrict__ eacc,
                                              long long* __restrict__ nquads) {
  double e = 0.0;
  long long nq = 0;
  int nbt = min(*nbonds, A.bond_cap);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nbt; t += gridDim.x * blockDim.x) {
    const ReboParams& P = T->rebo;
    int2 bd = bonds[t];
    int i = bd.x, j = bd.y;
    if (A.type[i] == T_C && A.type[j] == T_C) {
      D3 a_ = depos(A.x, A.y, A.z, i, j, bx);
      double rij = sqrt(norm2(a_));
      double wij, dwij;
      cutoff(rij, P.R[0][0][0], P.R[0][0][1], wij, dwij);
      D3 gi = d3(0, 0, 0), gj = d3(0, 0, 0);
      int nbi = A.bcnt[i], nbj = A.bcnt[j];
      for (int s = 0; s < nbi; s++) {
        BondRef k = bond_at(A, i, s);
        if (k.k == j) continue;
        double sin_ik = sqrt(norm2(cross(a_, k.rv))) / (rij * k.r);
        for (int u = 0; u < nbj; u++) {
          BondRef l = bond_at(A, j, u);
          if (l.k == i || l.k == k.k) continue;
          if (fabs(sin_ik) < 0.1) continue;
          double sin_jl = sqrt(norm2(cross(a_, l.rv))) / (rij * l.r);
          if (fabs(sin_jl) < 0.1) continue;
          nq++;
          int tk = A.type[k.k], tl = A.type[l.k];
          double wki = k.f, wjl = l.f;
          double zt = T->zetaT[tk][tl];
          D3 b_ = k.rv, c_ = l.rv;
          D3 Pp = cross(-a_, b_), Q = cross(a_, c_);
          double p2 = norm2(Pp), q2 = norm2(Q);
          double e2 = sqrt(p2) * sqrt(q2);
          double cs = (e2 == 0.0) ? 0.0 : dot(Pp, Q) / e2;
          double cd = cs;
          if (cd < -1.0) cd = -1.0;
          if (cd > 1.0) cd = 1.0;
          double c = -cd;
          double x = 0.5 * (1.0 + c);
          double vt = (256.0 / 405.0) * zt * pow(x, 5.0) - 0.1 * zt;
          D3 gk = d3(0, 0, 0), gl = d3(0, 0, 0);
          if (vt != 0.0) {
            double der = (256.0 / 405.0) * zt * pow(x, 5.0 - 1.0) * 0.5;
            if (!compat) der *= 5.0;
            double V = -1.0 * der * (wki * wij * wjl);
            if (V != 0.0 && e2 != 0.0) {
              D3 gP = Q / e2 - Pp * (cs / p2);
              D3 gQ = Pp / e2 - Q * (cs / q2);
              D3 da = cross(gP, b_) + cross(c_, gQ);
              D3 db = cross(a_, gP);
              D3 dc = cross(gQ, a_);
              gi = gi + (da + db) * V;
              gj = gj + (dc - da) * V;
              gk = gk - db * V;
              gl = gl - dc * V;
            }
          }
          // cutoff-weight derivatives
          if (k.df != 0.0) {
            D3 v = k.rv * (k.df * (wij * wjl * vt) / k.r);
            gi = gi + v;
            gk = gk - v;
          }
          if (dwij != 0.0) {
            D3 v = a_ * (dwij * (wki * wjl * vt) / rij);
            gi = gi + v;
            gj = gj - v;
          }
          if (l.df != 0.0) {
            D3 v = l.rv * (l.df * (wki * wij * vt) / l.r);
            gj = gj + v;
            gl = gl - v;
          }
          atomic_add3(grad, k.k, gk);
          atomic_add3(grad, l.k, gl);
          e += wki * wij * wjl * vt;
        }
      }
      atomic_add3(grad, i, gi);
      atomic_add3(grad, j, gj);
    }
  }
  block_accumulate(e, eacc);
  if (nq) atomicAdd((unsigned long long*)nquads, (unsigned long long)nq);
}

// ------------------------------------------------------------ ZBL
// FBZL::operator() (FBZL.cxx:70-127): ion-target and ion-ion pairs.
__device__ __forceinline__ bool zbl_target(int elem) {
  return elem == 1 || elem == 12 || elem == 64 || elem == 197 || elem == 108;
}
__device__ __forceinline__ bool zbl_ion(int elem) { return elem == 40 || elem == 131; }

__global__ void k_fbzl(Atoms A, Box bx, double R0, double R1, double* __restrict__ grad, double* __restrict__ eacc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < A.n && zbl_ion(A.elem[i])) {
    int cnt = A.nll_cnt[i];
    const int* list = A.nll + (size_t)i * A.cap_l;
    for (int q = 0; q < cnt; q++) {
      int j = list[q];
      bool jion = zbl_ion(A.elem[j]);
      if (!(jion || zbl_target(A.elem[j]))) continue;
      if (jion && j < i) continue;  // ion-ion pairs once
      int a = i < j ? i : j, b = i < j ? j : i;
      D3 rv = depos(A.x, A.y, A.z, a, b, bx);
      double d2 = norm2(rv);
      if (d2 > R1 * R1) continue;
      double r = sqrt(d2);
      double f, df;
      cutoff(r, R0, R1, f, df);
      if (r > R1) continue;
      double ZA = A.Z[a], ZB = A.Z[b];
      double AS = 8.8534e-1 * 0.53e-8 / (pow(ZA / kECharge, 0.23) + pow(ZB / kECharge, 0.23));
      double Y = r / AS;
      double e1 = exp(-3.1998 * Y), e2 = exp(-0.94229 * Y), e3 = exp(-0.4029 * Y), e4 = exp(-0.20162 * Y);
      double sum = 0.18175 * e1 + 0.50986 * e2 + 0.28022 * e3 + 0.02817 * e4;
      double val = ZA * ZB / r * sum;
      double der = -ZA * ZB / (r * r) * sum -
                   ZA * ZB / (r * AS) *
                       (0.18175 * 3.1998 * e1 + 0.50986 * 0.94229 * e2 + 0.28022 * 0.4029 * e3 + 0.02817 * 0.20162 * e4);
      double cr = der * f + df * val;
      D3 v = rv * (cr / r);
      atomic_add3(grad, a, v);
      atomic_add3(grad, b, -v);
      e += val * f;
    }
  }
  block_accumulate(e, eacc);
}

// ------------------------------------------------------------ Verlet
// eq.1 of SimLoop::executeMain (SimLoop.cxx:323-332): x += v dt + a dt^2/2,
// accumulated into both neighbour lists' displacements.
__global__ void k_drift(Atoms A, const double* __restrict__ dtp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n || A.fixed[i]) return;
  double dt = *dtp;
  double dx = A.vx[i] * dt + A.ax[i] * dt * dt / 2.0;
  double dy = A.vy[i] * dt + A.ay[i] * dt * dt / 2.0;
  double dz = A.vz[i] * dt + A.az[i] * dt * dt / 2.0;
  A.x[i] += dx;
  A.y[i] += dy;
  A.z[i] += dz;
  A.disp_s[3 * i] += dx;
  A.disp_s[3 * i + 1] += dy;
  A.disp_s[3 * i + 2] += dz;
  A.disp_l[3 * i] += dx;
  A.disp_l[3 * i + 1] += dy;
  A.disp_l[3 * i + 2] += dz;
}

__device__ __forceinline__ void wrap(double& c, double L, int& cnt) {
  if (L == kNoPbc) return;
  for (int q = 0; q < 64 && c < 0; q++) {
    c += L;
    --cnt;
  }
  for (int q = 0; q < 64 && c >= L; q++) {
    c -= L;
    ++cnt;
  }
}

// eq.2-4 (SimLoop.cxx:339-399), v_max over atoms with neighbours, the PBC
// wrap (:401, Atom.cxx:93-125), kinetic energy and net force partials.
__global__ void k_kick(Atoms A, Box bx, const double* __restrict__ dtp, double* __restrict__ red) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ek = 0.0, fx = 0.0, fy = 0.0, fz = 0.0, vm = 0.0;
  if (i < A.n) {
    double gx = A.grad[3 * i], gy = A.grad[3 * i + 1], gz = A.grad[3 * i + 2];
    fx = -gx;
    fy = -gy;
    fz = -gz;
    if (!A.fixed[i]) {
      double dt = *dtp;
      double M = A.mass[i];
      double hx = A.vx[i] + A.ax[i] * dt / 2.0;
      double hy = A.vy[i] + A.ay[i] * dt / 2.0;
      double hz = A.vz[i] + A.az[i] * dt / 2.0;
      double ax = fx / M, ay = fy / M, az = fz / M;
      A.ax[i] = ax;
      A.ay[i] = ay;
      A.az[i] = az;
      double vx = hx + ax * dt / 2.0, vy = hy + ay * dt / 2.0, vz = hz + az * dt / 2.0;
      A.vx[i] = vx;
      A.vy[i] = vy;
      A.vz[i] = vz;
      double v = sqrt(vx * vx + vy * vy + vz * vz);
      if (A.nls_cnt[i] > 0 || A.nll_cnt[i] > 0) vm = v;
    }
    double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i];
    double vv = sqrt(vx * vx + vy * vy + vz * vz);
    ek = A.mass[i] * (vv * vv) / 2.0;
    wrap(A.x[i], bx.L[0], A.pcx[i]);
    wrap(A.y[i], bx.L[1], A.pcy[i]);
    wrap(A.z[i], bx.L[2], A.pcz[i]);
  }
  block_accumulate(ek, red + 0);
  block_accumulate(fx, red + 1);
  block_accumulate(fy, red + 2);
  block_accumulate(fz, red + 3);
  // v_max: max of non-negative doubles via integer max on the bit pattern
  __shared__ unsigned long long vmx;
  if (threadIdx.x == 0) vmx = 0ull;
  __syncthreads();
  atomicMax(&vmx, (unsigned long long)__double_as_longlong(vm));
  __syncthreads();
  if (threadIdx.x == 0) atomicMax((unsigned long long*)(red + 4), vmx);
}

// First force evaluation (iteration 0, SimLoop.cxx:305-321): a = -grad/M.
__global__ void k_init_acc(Atoms A) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n || A.fixed[i]) return;
  double M = A.mass[i];
  A.ax[i] = -A.grad[3 * i] / M;
  A.ay[i] = -A.grad[3 * i + 1] / M;
  A.az[i] = -A.grad[3 * i + 2] / M;
}

// Per-step bookkeeping: energies out, dt rule (SimLoop.cxx:409-416), reset.
__global__ void k_step_end(int nterms, double* __restrict__ eacc, double* __restrict__ red, double* __restrict__ dtp,
                           int fixed_dt, double* __restrict__ out, int step) {
  if (threadIdx.x != 0) return;
  double* o = out + (size_t)step * 8;
  double ep = 0;
  for (int t = 0; t < nterms; t++) ep += eacc[t];
  o[0] = ep;
  o[1] = red[0];
  o[2] = *dtp;
  o[3] = sqrt(red[1] * red[1] + red[2] * red[2] + red[3] * red[3]);
  if (!fixed_dt) {
    double vmax = red[4];
    const double dt_max = 5e-16, dt_min = 1e-20;
    double dt;
    if (vmax != 0.0) dt = 0.05 * (1.0 * kAo) / vmax;
    else dt = dt_max;
    if (dt > dt_max) dt = dt_max;
    if (dt < dt_min) dt = dt_min;
    *dtp = dt;
  }
  for (int q = 0; q < 8; q++) red[q] = 0.0;
}

__global__ void k_zero(double* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void k_zero_i(int* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace mdg
