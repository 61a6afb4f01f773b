// Neighbour lists, skin test, ZBL, the velocity-Verlet kernels and the
// per-system reductions (FP64).  A context holds one system or a batch of
// independent systems; A.sys[i] is atom i's system, A.sys_off the atom ranges.
// Kernels are compiled without FMA contraction so that a*b+c rounds twice, as
// in the reference's x86-64 build.
#include "mdg_kernels.hpp"

namespace mdg {

// The packed copy of atom i at (x, y, z): reduced into the box on the
// periodic axes on which the atom itself is open (packed_coord), else as is.
__device__ __forceinline__ double4 packed_pos(const Atoms& A, int i, double x, double y, double z) {
  if (A.open_axes) {
    const unsigned m = A.open_axes[i];
    if (m) {
      const double* L = A.boxL + 3 * A.sys[i];
      if (m & 1u) x = packed_coord(x, L[0]);
      if (m & 2u) y = packed_coord(y, L[1]);
      if (m & 4u) z = packed_coord(z, L[2]);
    }
  }
  return make_double4(x, y, z, double(A.type[i]));
}


// ----------------------------------------------------------------- helpers

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block (256 threads) sum; the result is valid in thread 0.
__device__ double block_sum(double v) {
  __shared__ double red[16];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) s += red[q];
  }
  return s;
}

// ------------------------------------------------------------ cell lists
// Every system has its own grid; cell ids are offset by cell_off[system].
// Per-system cell passes: system s owns the cells [cell_off[s], cell_off[s+1])
// and the slots [sys_off[s], sys_off[s+1]) of cell_atoms, so only the systems
// whose rebuild flag is set (sflags[2*s + list]) are redone.
__global__ void k_cell_clear(const int* __restrict__ flag, const int* __restrict__ sflags, int list_id,
                             const int* __restrict__ cell_off, int* __restrict__ cnt) {
  if (!*flag) return;
  int s = blockIdx.x;
  if (!sflags[2 * s + list_id]) return;
  for (int c = cell_off[s] + threadIdx.x; c < cell_off[s + 1]; c += blockDim.x) cnt[c] = 0;
}

__global__ void k_cell_assign(const int* __restrict__ flag, const int* __restrict__ sflags, int list_id, int n,
                              const double4* __restrict__ pos, const CellGrid* __restrict__ grids,
                              const int* __restrict__ cell_off, const int* __restrict__ sys,
                              int* __restrict__ cell_of, int* __restrict__ cell_cnt) {
  if (!*flag) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = sys[i];
  if (!sflags[2 * s + list_id]) return;
  const CellGrid& g = grids[s];
  int c[3];
  double4 pi = pos[i];
  double p[3] = {pi.x, pi.y, pi.z};
  for (int d = 0; d < 3; d++) {
    int q = (int)floor((p[d] - g.lo[d]) / g.w[d]);
    if (q < 0) q = 0;
    if (q >= g.n[d]) q = g.n[d] - 1;
    c[d] = q;
  }
  int id = cell_off[s] + (c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
  cell_of[i] = id;
  atomicAdd(&cell_cnt[id], 1);
}

// Exclusive scan of one system's cell counts (one 1024-thread workgroup per
// system), offset by the system's first atom slot.
__global__ void k_cell_scan(const int* __restrict__ flag, const int* __restrict__ sflags, int list_id,
                            const int* __restrict__ cell_off, const int* __restrict__ sys_off, int* __restrict__ cnt,
                            int* __restrict__ start) {
  if (!*flag) return;
  int sy = blockIdx.x;
  if (!sflags[2 * sy + list_id]) return;
  __shared__ int part[1024];
  int c0 = cell_off[sy], ncell = cell_off[sy + 1] - c0;
  int t = threadIdx.x;
  int chunk = (ncell + 1023) / 1024;
  int b = c0 + min(t * chunk, ncell), e = c0 + min((t + 1) * chunk, ncell);
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
  int run = sys_off[sy] + part[t] - s;
  for (int q = b; q < e; q++) {
    start[q] = run;
    run += cnt[q];
    cnt[q] = 0;  // reused as fill cursor
  }
  if (t == 1023) start[c0 + ncell] = sys_off[sy] + part[1023];
}

__global__ void k_cell_fill(const int* __restrict__ flag, const int* __restrict__ sflags, int list_id, int n,
                            const int* __restrict__ sys, const int* __restrict__ cell_of,
                            const int* __restrict__ start, int* __restrict__ cursor, int* __restrict__ cell_atoms) {
  if (!*flag) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!sflags[2 * sys[i] + list_id]) return;
  int c = cell_of[i];
  int p = atomicAdd(&cursor[c], 1);
  cell_atoms[start[c] + p] = i;
}

// Neighbour list build, one wavefront per atom (NeighbourList.cxx:31-106):
// pair (i,j) is listed iff both atoms are handled and d^2 < ((1+0.44) Rc)^2
// with d^2 = dx*dx+dy*dy+dz*dz of the minimum-image displacement.  The
// candidates of all stencil cells are swept 64 per round (a prefix over the
// cells' counts maps a lane to its cell), compacted with a ballot/popcount
// prefix into LDS, then (SORT) sorted ascending, the reference's list order,
// with a bitonic sort, or (!SORT, the long list of the mixed form, whose
// users need only the split) written as the j < i entries, then the j > i
// entries, each in candidate order.  The FP64 form keeps every list sorted,
// so its sums run in an order that no rebuild schedule changes (a run
// resumed from a saved state repeats the uninterrupted one bit for bit).
// Only atoms whose system asked for a rebuild (sflags[2*s+list]) are redone;
// the grid strides over the atoms, so a launch with nothing to rebuild costs
// a few thousand empty waves, not one per atom.
template <int CAP, bool SORT, bool APBC>
__global__ void __launch_bounds__(256) k_nl_build(const int* __restrict__ flag, const int* __restrict__ sflags,
                                                  int list_id, int n, const double4* __restrict__ pos,
                                                  const unsigned char* __restrict__ handled, Box bx,
                                                  const CellGrid* __restrict__ grids,
                                                  const int* __restrict__ cell_off, double range2,
                                                  const int* __restrict__ cell_of, const int* __restrict__ start,
                                                  const int* __restrict__ cell_atoms, int* __restrict__ cnt,
                                                  int* __restrict__ list, double* __restrict__ disp, unsigned* err,
                                                  int* __restrict__ split, const double* __restrict__ xs,
                                                  const double* __restrict__ ys, const double* __restrict__ zs) {
  if (!*flag) return;
  __shared__ int buf[4][CAP];
  __shared__ int s_pre[4][32], s_st[4][32];  // stencil cells: first candidate, first atom
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = blockIdx.x * 4 + w; i < n; i += gridDim.x * 4) {
    int sy = bx.sys[i];
    if (!sflags[2 * sy + list_id]) continue;  // wave-uniform
    // APBC: the systems whose atoms carry their own PBC vectors, and only those
    if (APBC != (bx.apbc != nullptr && bx.apbc[sy] != 0)) continue;
    const CellGrid& g = grids[sy];
    const Frame fr = frame_of(bx, i);
    int base_cell = cell_off[sy];
    int* my = buf[w];
    if (lane == 0) {
      disp[3 * i] = 0.0;
      disp[3 * i + 1] = 0.0;
      disp[3 * i + 2] = 0.0;
    }
    int m = 0;
    if (handled[i]) {
      double4 pi = pos[i];
      int ci = cell_of[i] - base_cell;
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
      // the stencil's cells (at most 27), one per lane: start and atom count,
      // then an exclusive prefix over the counts, so that the candidates of
      // all cells are swept 64 at a time (the list is sorted afterwards, so
      // the sweep order does not matter)
      int e1 = hi[1] - lo[1] + 1, e2 = hi[2] - lo[2] + 1;
      int ncs = (hi[0] - lo[0] + 1) * e1 * e2;
      int cst = 0, ccn = 0;
      if (lane < ncs) {
        int q0 = cc[0] + lo[0] + lane / (e1 * e2), q1 = cc[1] + lo[1] + (lane / e2) % e1, q2 = cc[2] + lo[2] + lane % e2;
        bool in = true;
        if (g.periodic[0]) q0 = (q0 + g.n[0]) % g.n[0];
        else in = in && q0 >= 0 && q0 < g.n[0];
        if (g.periodic[1]) q1 = (q1 + g.n[1]) % g.n[1];
        else in = in && q1 >= 0 && q1 < g.n[1];
        if (g.periodic[2]) q2 = (q2 + g.n[2]) % g.n[2];
        else in = in && q2 >= 0 && q2 < g.n[2];
        if (in) {
          int cid = base_cell + (q0 * g.n[1] + q1) * g.n[2] + q2;
          cst = start[cid];
          ccn = start[cid + 1] - cst;
        }
      }
      int incl = ccn;
      for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      int total = __shfl(incl, 63, 64);
      if (lane < 32) {
        s_pre[w][lane] = incl - ccn;
        s_st[w][lane] = cst;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (int base = 0; base < total; base += 64) {
        int t = base + lane;
        bool ok = false;
        int j = -1;
        if (t < total) {
          int lo_c = 0, hi_c = ncs - 1;  // last stencil cell whose first candidate is <= t
          while (lo_c < hi_c) {
            int mid = (lo_c + hi_c + 1) >> 1;
            if (s_pre[w][mid] <= t) lo_c = mid;
            else hi_c = mid - 1;
          }
          j = cell_atoms[s_st[w][lo_c] + (t - s_pre[w][lo_c])];
          if (j != i && handled[j]) {
            double4 pj = pos[j];
            D3 d;
            if constexpr (APBC) d = depos_atoms(bx, xs, ys, zs, i, j);  // each atom's own PBC
            else d = mdg::disp(bx, fr, i, j, pi, pj);  // qualified: the argument disp names the displacements
            double d2 = d.x * d.x + d.y * d.y + d.z * d.z;
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
    if (m > CAP) {
      if (lane == 0) raise(sys_err(err, sy), CAP >= 256 ? ERR_LONG_NL_OVERFLOW : ERR_SHORT_NL_OVERFLOW);
      m = CAP;
    }
    if constexpr (!SORT) {  // j < i entries first, then j > i, each in candidate order
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      int lt = 0;
      for (int q = lane; q < m; q += 64) lt += (my[q] < i);
      lt = wave_sum(lt);
      int* row = list + (size_t)i * CAP;
      const unsigned long long lanes_below = (1ull << lane) - 1ull;
      for (int b = 0, a0 = 0, a1 = lt; b < m; b += 64) {
        const int q = b + lane;
        const int v = (q < m) ? my[q] : 0;
        const bool low = q < m && v < i, high = q < m && v > i;
        const unsigned long long ml = __ballot(low), mh = __ballot(high);
        if (low) row[a0 + __popcll(ml & lanes_below)] = v;
        if (high) row[a1 + __popcll(mh & lanes_below)] = v;
        a0 += __popcll(ml);
        a1 += __popcll(mh);
      }
      if (lane == 0) {
        cnt[i] = m;
        if (split) split[i] = lt;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      continue;
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
    for (int q = lane; q < m; q += 64) list[(size_t)i * CAP + q] = my[q];
    if (lane == 0) cnt[i] = m;
    if (split) {  // entries below i: the row's j > i part starts there
      int lt = 0;
      for (int q = lane; q < m; q += 64) lt += (my[q] < i);
      lt = wave_sum(lt);
      if (lane == 0) split[i] = lt;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

#define MDG_NL_BUILD(CAP, SORT, APBC)                                                                          \
  template __global__ void k_nl_build<CAP, SORT, APBC>(                                                      \
      const int*, const int*, int, int, const double4*, const unsigned char*, Box, const CellGrid*, const int*, \
      double, const int*, const int*, const int*, int*, int*, double*, unsigned*, int*, const double*,         \
      const double*, const double*);
MDG_NL_BUILD(64, true, false)
MDG_NL_BUILD(512, true, false)
MDG_NL_BUILD(512, false, false)
MDG_NL_BUILD(64, true, true)
MDG_NL_BUILD(512, true, true)
MDG_NL_BUILD(512, false, true)

// ------------------------------------------------------------ skin test
// Top-2 accumulated displacement per system and list (NeighbourList.cxx:
// 108-133): rebuild iff d1 + d2 > 0.44 Rc.  The two largest values of a set
// are exact whatever the merge order.
__device__ __forceinline__ void top2_insert(double v, double& d1, double& d2) {
  d2 = fmax(d2, fmin(v, d1));  // v >= d1: d1 moves down; else v competes for second place
  d1 = fmax(d1, v);
}

__device__ __forceinline__ void top2_merge(double& d1, double& d2, double e1, double e2) {
  top2_insert(e1, d1, d2);
  top2_insert(e2, d1, d2);
}

__device__ __forceinline__ void top2_wave(double& d1, double& d2) {
  for (int o = 32; o > 0; o >>= 1) {
    double e1 = __shfl_xor(d1, o, 64), e2 = __shfl_xor(d2, o, 64);
    top2_merge(d1, d2, e1, e2);
  }
}

// Per-system reductions run in two stages: a 256-thread block per tile of
// at most kRedTile atoms of one system (rb[tile] = {system, first atom})
// writes its partial results to part[8 * tile + k]; then a wave per system
// merges that system's tiles, rb_off[s] .. rb_off[s+1], in tile order, so a
// system's result does not depend on the other systems of the batch.
__device__ __forceinline__ int tile_end(const int2 t, const int* __restrict__ sys_off) {
  return min(t.y + kRedTile, sys_off[t.x + 1]);
}

// The two largest displacement lengths of each list over one tile.
// The two largest values per list over a 256-thread block -- v holds this
// thread's (d1, d2) of the short list, the long list, the LJ pair rows and
// the metal rows -- written to part[8 * blockIdx.x ..].
__device__ __forceinline__ void top2_tile_out(double (&v)[8], double* __restrict__ part) {
  __shared__ double sh[4][8];
#pragma unroll
  for (int k = 0; k < 4; k++) top2_wave(v[2 * k], v[2 * k + 1]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 8; k++) sh[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < (int)(blockDim.x >> 6); q++)
#pragma unroll
      for (int k = 0; k < 4; k++) top2_merge(x[2 * k], x[2 * k + 1], sh[q][2 * k], sh[q][2 * k + 1]);
    double* o = part + 8 * (size_t)blockIdx.x;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = x[k];
  }
}

__global__ void __launch_bounds__(256) k_sys_disp_part(const int2* __restrict__ rb, const int* __restrict__ sys_off,
                                                       const double* __restrict__ ds, const double* __restrict__ dl,
                                                       const double* __restrict__ dp, const double* __restrict__ de,
                                                       double* __restrict__ part) {
  const int2 t = rb[blockIdx.x];
  const int b = t.y, e = tile_end(t, sys_off);
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto len = [](const double* d, int i) {
    return sqrt(d[3 * i] * d[3 * i] + d[3 * i + 1] * d[3 * i + 1] + d[3 * i + 2] * d[3 * i + 2]);
  };
  for (int i = b + threadIdx.x; i < e; i += blockDim.x) {
    top2_insert(len(ds, i), v[0], v[1]);
    top2_insert(len(dl, i), v[2], v[3]);
    if (dp) top2_insert(len(dp, i), v[4], v[5]);
    if (de) top2_insert(len(de, i), v[6], v[7]);
  }
  top2_tile_out(v, part);
}

// One wave per system, the skin test of NeighbourList::MovedTooMuch (the
// two largest displacements): sflags[2*s+{0,1}], pflags[s] and the
// any-system flags[0..2].  The LJ pair rows are redone with the long list or when their
// own displacements pass the margin (only when have_p).
__global__ void __launch_bounds__(64) k_sys_disp_merge(const int* __restrict__ rb_off, const double* __restrict__ part,
                                                       double lim_s, double lim_l, double lim_p, int have_p,
                                                       double lim_e, int have_e, int* __restrict__ sflags,
                                                       int* __restrict__ pflags, int* __restrict__ eflags,
                                                       int* __restrict__ flags, long long* __restrict__ rebuilds) {
  const int s = blockIdx.x;
  double x1 = 0, x2 = 0, y1 = 0, y2 = 0, z1 = 0, z2 = 0, u1 = 0, u2 = 0;
  for (int q = rb_off[s] + threadIdx.x; q < rb_off[s + 1]; q += 64) {
    const double* o = part + 8 * (size_t)q;
    top2_merge(x1, x2, o[0], o[1]);
    top2_merge(y1, y2, o[2], o[3]);
    top2_merge(z1, z2, o[4], o[5]);
    top2_merge(u1, u2, o[6], o[7]);
  }
  top2_wave(x1, x2);
  top2_wave(y1, y2);
  top2_wave(z1, z2);
  top2_wave(u1, u2);
  if (threadIdx.x == 0) {
    int fs = (x1 + x2 > lim_s) ? 1 : 0;
    int fl = (y1 + y2 > lim_l) ? 1 : 0;
    int fp = (have_p && (fl || z1 + z2 > lim_p)) ? 1 : 0;
    sflags[2 * s] |= fs;
    sflags[2 * s + 1] |= fl;
    if (fs) {
      atomicOr(&flags[0], 1);
      atomicAdd((unsigned long long*)&rebuilds[0], 1ull);
    }
    if (fl) {
      atomicOr(&flags[1], 1);
      atomicAdd((unsigned long long*)&rebuilds[1], 1ull);
    }
    if (fp) {
      pflags[s] = 1;
      atomicOr(&flags[2], 1);
    }
    if (have_e && (fl || u1 + u2 > lim_e)) {
      eflags[s] = 1;
      atomicOr(&flags[3], 1);
    }
  }
}

__global__ void k_clear_flags(int nsys, int* __restrict__ flags, int* __restrict__ sflags, int* __restrict__ pflags,
                              int* __restrict__ eflags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 4) flags[i] = 0;
  if (i < 2 * nsys) sflags[i] = 0;
  if (i < nsys) pflags[i] = 0;
  if (eflags && i < nsys) eflags[i] = 0;
}

// The metal rows of the flagged systems (their skin test restarts: disp_e
// of a rebuilt atom is zeroed).  A row keeps the long row's order.
__global__ void __launch_bounds__(256) k_metal_rows(Atoms A, Box bx, const int* __restrict__ flag,
                                                    const int* __restrict__ eflags, double rcand) {
  if (!*flag) return;
  constexpr int LW = 16;
  const int lane = threadIdx.x & 63;
  const int sl = lane % LW;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mine = ((1ull << LW) - 1ull) << (lane / LW * LW);
  const int i = blockIdx.x * (blockDim.x / LW) + threadIdx.x / LW;
  if (i >= A.n || !eflags[A.sys[i]]) return;  // group-uniform
  if (sl == 0) {
    A.disp_e[3 * i] = 0.0;
    A.disp_e[3 * i + 1] = 0.0;
    A.disp_e[3 * i + 2] = 0.0;
  }
  const double4 pi = A.p[i];
  if ((int)pi.w != T_OTHER) {
    if (sl == 0) A.ml_cnt[i] = 0;
    return;
  }
  const int cnt = A.nll_cnt[i];
  const int* L = A.nll + (size_t)i * A.cap_l;
  int* my = A.ml + (size_t)i * kMetalRow;
  const Frame fr = frame_of(bx, i);
  const double rc2 = rcand * rcand;
  int m = 0;
  for (int b0 = 0; b0 < cnt; b0 += LW) {  // group-uniform trip count
    const int q = b0 + sl;
    bool in = false;
    int j = 0;
    if (q < cnt) {
      j = L[q];
      const double4 pj = A.p[j];
      if ((int)pj.w == T_OTHER) {
        const D3 dd = disp(bx, fr, i, j, pi, pj);
        const double dx = dd.x, dy = dd.y, dz = dd.z;
        in = dx * dx + dy * dy + dz * dz < rc2;
      }
    }
    const unsigned long long msk = __ballot(in) & mine;
    if (in) {
      const int slot = m + __popcll(msk & below);
      if (slot < kMetalRow) my[slot] = j;
    }
    m += __popcll(msk);
  }
  if (sl == 0) A.ml_cnt[i] = (m <= kMetalRow) ? m : -1;  // -1: the density pass sweeps the long row
}

// Bonds (i<j) in the order of i's list, appended to one array; one atomic
// per workgroup (blockDim 256) reserves the group's bonds, placed by an
// inclusive scan of the per-atom counts.
__global__ void __launch_bounds__(256) k_bond_enum(Atoms A, int2* __restrict__ out, int* __restrict__ count, int cap,
                                                   unsigned* err, long long* __restrict__ ntrip) {
  __shared__ int s_w[4];
  __shared__ int s_base;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int nb = (i < A.n) ? A.bcnt[i] : 0;
  // sigma-pi triplets (bond, side, neighbour) of this call's REBO bonds: a
  // centre with nb bonds takes nb - 1 neighbours for each of its nb bonds
  count_add(ntrip, (long long)nb * (nb > 0 ? nb - 1 : 0));
  int mine = 0;
  for (int s = 0; s < nb; s++) mine += (A.bj[i * A.maxb + s] > i);
  int incl = mine;
  for (int o = 1; o < 64; o <<= 1) {
    int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    s_base = total > 0 ? atomicAdd(count, total) : 0;
  }
  __syncthreads();
  int base = s_base + incl - mine;
  for (int u = 0; u < w; u++) base += s_w[u];
  if (!mine) return;
  for (int s = 0; s < nb; s++) {
    int j = A.bj[i * A.maxb + s];
    if (j > i) {
      if (base < cap) out[base] = make_int2(i, j);
      else raise(err, ERR_QUEUE_OVERFLOW);
      base++;
    }
  }
}

// ------------------------------------------------------------ ZBL
// FBZL::operator() (FBZL.cxx:70-127): ion-target and ion-ion pairs.
__device__ __forceinline__ bool zbl_target(int elem) {
  return elem == 1 || elem == 12 || elem == 64 || elem == 197 || elem == 108;
}
__device__ __forceinline__ bool zbl_ion(int elem) { return elem == 40 || elem == 131; }

// One wavefront per ion (ions[w]): its lanes share the ion's long-list row;
// the pair energies go to the ion's slot, one sum per wave.
template <bool VIR>
__global__ void __launch_bounds__(256) k_fbzl(Atoms A, Box bx, const int* __restrict__ ions, int nions, double R0,
                                              double R1, double* __restrict__ grad, double* __restrict__ eatom,
                                              long long* __restrict__ vir) {
  const int w = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= nions) return;  // wave-uniform
  const int i = ions[w];
  Virial<VIR> vw(vir, 1.0);
  if constexpr (VIR) vw.at(A.sys[i]);
  double e = 0.0;
  const int cnt = A.nll_cnt[i];
  const int* list = A.nll + (size_t)i * A.cap_l;
  for (int q = lane; q < cnt; q += 64) {
    int j = list[q];
    bool jion = zbl_ion(A.elem[j]);
    if (!(jion || zbl_target(A.elem[j]))) continue;
    if (jion && j < i) continue;  // ion-ion pairs once
    int a = i < j ? i : j, b = i < j ? j : i;
    D3 rv = depos(A.p, a, b, bx);
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
    atomic_add3(GradSink{grad, A.gfix}, a, v);
    atomic_add3(GradSink{grad, A.gfix}, b, -v);
    vw.push(rv, v);
    e += val * f;
  }
  e = wave_sum(e);
  if (lane == 0) energy_add(A, eatom, i, e);
  vw.flush_wave();
}
template __global__ void k_fbzl<false>(Atoms, Box, const int*, int, double, double, double*, double*, long long*);
template __global__ void k_fbzl<true>(Atoms, Box, const int*, int, double, double, double*, double*, long long*);

// FLJ::operator() (FLJ.cxx:256-341) for the pairs {H,C} x {Cu,Au}: each metal
// atom takes its hydrocarbon partners from the long list (so every pair once);
// ZBL below x[0], the cubic bridge to x[1], LJ times the cosine cutoff beyond.
// The pair's energy goes to its lower-index atom, as the reference's i < j
// loop sums it.
__device__ __forceinline__ int flj_index(int elem) {
  switch (elem) {
    case 64: return 0;   // Cu
    case 1: return 1;    // H
    case 12: return 2;   // C
    case 197: return 3;  // Au
    default: return -1;
  }
}

template <bool VIR>
__global__ void k_flj(Atoms A, Box bx, const FLJTable* __restrict__ T, double* __restrict__ grad,
                      double* __restrict__ eatom, long long* __restrict__ vir) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= A.n) return;
  const int em = A.elem[m];
  if (em != 64 && em != 197) return;
  Virial<VIR> vw(vir, 1.0);
  if constexpr (VIR) vw.at(A.sys[m]);
  const int im = flj_index(em);
  const int cnt = A.nll_cnt[m];
  const int* list = A.nll + (size_t)m * A.cap_l;
  const double R1 = T->R1;
  for (int q = 0; q < cnt; q++) {
    const int h = list[q];
    const int eh = A.elem[h];
    if (eh != 1 && eh != 12) continue;
    const int a = m < h ? m : h, b = m < h ? h : m;  // AtomsPair(atom_i, atom_j), i < j
    const int ia = (a == m) ? im : flj_index(eh), ib = (a == m) ? flj_index(eh) : im;
    D3 rv = depos(A.p, a, b, bx);
    double d2 = norm2(rv);
    if (d2 > R1 * R1) continue;  // FPairwise::probablyAreNeighbours
    double r = sqrt(d2);
    double val, der;
    if (r < T->x[0]) {
      double ZA = A.Z[a], ZB = A.Z[b];
      double AS = 8.8534e-1 * 0.53e-8 / (pow(ZA / kECharge, 0.23) + pow(ZB / kECharge, 0.23));
      double Y = r / AS;
      double e1 = exp(-3.1998 * Y), e2 = exp(-0.94229 * Y), e3 = exp(-0.4029 * Y), e4 = exp(-0.20162 * Y);
      double sum = 0.18175 * e1 + 0.50986 * e2 + 0.28022 * e3 + 0.02817 * e4;
      val = ZA * ZB / r * sum;
      der = -ZA * ZB / (r * r) * sum -
            ZA * ZB / (r * AS) *
                (0.18175 * 3.1998 * e1 + 0.50986 * 0.94229 * e2 + 0.28022 * 0.4029 * e3 + 0.02817 * 0.20162 * e4);
    } else if (r < T->x[1]) {  // the cubic through both ends' values and slopes (Spline.cxx:72-98)
      const double* v = T->v[ia][ib];
      const double* dd = T->d[ia][ib];
      double hh = T->x[1] - T->x[0], t = (r - T->x[0]) / hh;
      val = (1 + 2 * t) * (1 - t) * (1 - t) * v[0] + t * (1 - t) * (1 - t) * hh * dd[0] + t * t * (3 - 2 * t) * v[1] +
            t * t * (t - 1) * hh * dd[1];
      der = (6 * t * (t - 1) * v[0] - 6 * t * (t - 1) * v[1]) / hh + (1 - t) * (1 - 3 * t) * dd[0] +
            t * (3 * t - 2) * dd[1];
    } else {
      double s = T->sigma[ia][ib] / r, s6 = s;
      for (int u = 2; u <= 6; u++) s6 *= s;
      double s12 = s6 * s6;
      double lv = 4.0 * T->eps[ia][ib] * (s12 - s6);
      double ld = 4.0 * T->eps[ia][ib] * (-12.0 * s12 / r + 6.0 * s6 / r);
      double f, df;
      cutoff(r, T->R0, T->R1, f, df);
      val = f * lv;
      der = ld * f + lv * df;
    }
    D3 v = rv * (der / r);
    atomic_add3(GradSink{grad, A.gfix}, a, v);
    atomic_add3(GradSink{grad, A.gfix}, b, -v);
    vw.push(rv, v);
    energy_add(A, eatom, a, val);
  }
  vw.flush();
}
template __global__ void k_flj<false>(Atoms, Box, const FLJTable*, double*, double*, long long*);
template __global__ void k_flj<true>(Atoms, Box, const FLJTable*, double*, double*, long long*);

// ------------------------------------------------------------ Ackland EAM
// Ackland::operator() (Ackland.cxx:31-57): E = sum_i F(rho_i) + sum_{i<j} V,
// F = -sqrt(rho), rho_i = sum_j g(r_ij).  Two passes over the long list with
// 16 lanes per metal atom: the density pass stores F'(rho_i) (0 where rho is
// 0, as F() applies no gradient there, :210-221) beside the atom's element
// and F_i; the force pass
// gathers, per ordered pair, dV/dr + (F'_i + F'_j) dg/dr along x_i - x_j into
// atom i alone (no scatter to j), and V of the pairs j > i into i's energy.
__device__ __forceinline__ int eam_index(int elem) {
  switch (elem) {
    case 64: return 0;   // Cu
    case 108: return 1;  // Ag
    case 197: return 2;  // Au
    default: return -1;
  }
}

// The element's density knot sum phi_e(r) and its slope (PhiCap/dPhiCap,
// Ackland.cxx:131-166).
__device__ __forceinline__ void eam_phi(const EAMTable* __restrict__ T, int e, double r, double& p, double& dp) {
  p = 0.0;
  dp = 0.0;
  if (r >= T->Rk[e][0]) return;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const double rt = T->Rk[e][k] - r;
    if (rt <= 0) continue;
    p += T->Ak[e][k] * rt * rt * rt;
    dp += T->Ak[e][k] * 3.0 * rt * rt * (-1.0);
  }
}

// g of the pair and dg/dr (Ackland.cxx:168-208): the like phi, or the
// geometric mean of both like phi, whose slope is applied only inside both
// elements' ranges.
__device__ __forceinline__ void eam_g(const EAMTable* __restrict__ T, int a, int b, double r, double& g, double& dg) {
  if (a == b) {
    eam_phi(T, a, r, g, dg);
    return;
  }
  double p1, d1, p2, d2;
  eam_phi(T, a, r, p1, d1);
  eam_phi(T, b, r, p2, d2);
  g = sqrt(p1 * p2);
  dg = (r < T->Rk[a][0] && r < T->Rk[b][0]) ? 0.5 / g * (d1 * p2 + p1 * d2) : 0.0;
}

// V of the pair and dV/dr (Phi, Ackland.cxx:59-129).
__device__ __forceinline__ void eam_pair(const EAMTable* __restrict__ T, int a, int b, double ZA, double ZB, double r,
                                         double& v, double& dv) {
  v = 0.0;
  dv = 0.0;
  if (r >= T->rk[a][b][1]) return;
  if (r < T->x[0]) {
    const double AS = T->AS[a][b], Y = r / AS;
    const double e1 = exp(-3.1998 * Y), e2 = exp(-0.94229 * Y), e3 = exp(-0.4029 * Y), e4 = exp(-0.20162 * Y);
    const double sum = 0.18175 * e1 + 0.50986 * e2 + 0.28022 * e3 + 0.02817 * e4;
    v = ZA * ZB / r * sum;
    dv = -ZA * ZB / (r * r) * sum -
         ZA * ZB / (r * AS) * (0.18175 * 3.1998 * e1 + 0.50986 * 0.94229 * e2 + 0.28022 * 0.4029 * e3 + 0.02817 * 0.20162 * e4);
  } else if (r < T->x[1]) {  // the cubic through both ends' values and slopes (Spline.cxx:72-98)
    const double* vv = T->v[a][b];
    const double* dd = T->d[a][b];
    const double hh = T->x[1] - T->x[0], t = (r - T->x[0]) / hh;
    v = (1 + 2 * t) * (1 - t) * (1 - t) * vv[0] + t * (1 - t) * (1 - t) * hh * dd[0] + t * t * (3 - 2 * t) * vv[1] +
        t * t * (t - 1) * hh * dd[1];
    dv = (6 * t * (t - 1) * vv[0] - 6 * t * (t - 1) * vv[1]) / hh + (1 - t) * (1 - 3 * t) * dd[0] + t * (3 * t - 2) * dd[1];
  } else {
    const int nk = (a == b) ? 6 : 3;
    for (int k = 1; k <= nk; k++) {
      const double rt = T->rk[a][b][k] - r;
      if (rt <= 0) continue;
      v += T->ak[a][b][k] * rt * rt * rt;
      dv += T->ak[a][b][k] * 3.0 * rt * rt * (-1.0);
    }
  }
}

__device__ __forceinline__ double AcklandModel::rmax(const Table* T) { return T->rmax; }
__device__ __forceinline__ void AcklandModel::g(const Table* T, int a, int b, double r, double& g, double& dg) {
  eam_g(T, a, b, r, g, dg);
}
__device__ __forceinline__ void AcklandModel::pair(const Table* T, int a, int b, double r, double& v, double& dv) {
  eam_pair(T, a, b, T->Z[a], T->Z[b], r, v, dv);
}
__device__ __forceinline__ void AcklandModel::embed(const Table*, double rho, double& F, double& Fp) {
  F = -sqrt(rho);
  Fp = (rho != 0.0) ? 0.5 / F : 0.0;
}

__device__ __forceinline__ double TBModel::rmax(const Table* T) { return T->R1; }
// g = f(r) exp(-beta r) (TightBinding.cxx:105-130)
__device__ __forceinline__ void TBModel::g(const Table* T, int, int, double r, double& g, double& dg) {
  double f, df;
  cutoff(r, T->R0, T->R1, f, df);
  const double e = exp(-T->beta * r);
  g = f * e;
  dg = e * (-T->beta) * f + df * e;
}
// V (TightBinding.cxx:59-102): 0 where f is 0; Born-Mayer, the bridge, then
// f Phi0 exp(-alpha r)
__device__ __forceinline__ void TBModel::pair(const Table* T, int, int, double r, double& v, double& dv) {
  double f, df;
  cutoff(r, T->R0, T->R1, f, df);
  v = 0.0;
  dv = 0.0;
  if (f == 0.0) return;
  if (r < T->x[0]) {
    v = T->A * exp(-T->B * r);
    dv = -T->B * T->A * exp(-T->B * r);
  } else if (r < T->x[1]) {  // the cubic through both ends' values and slopes (Spline.cxx:72-98)
    const double hh = T->x[1] - T->x[0], t = (r - T->x[0]) / hh;
    v = (1 + 2 * t) * (1 - t) * (1 - t) * T->v[0] + t * (1 - t) * (1 - t) * hh * T->d[0] +
        t * t * (3 - 2 * t) * T->v[1] + t * t * (t - 1) * hh * T->d[1];
    dv = (6 * t * (t - 1) * T->v[0] - 6 * t * (t - 1) * T->v[1]) / hh + (1 - t) * (1 - 3 * t) * T->d[0] +
         t * (3 * t - 2) * T->d[1];
  } else {
    const double e = T->Phi0 * exp(-T->alpha * r);
    v = f * e;
    dv = e * (-T->alpha) * f + df * e;
  }
}
// F = -c sqrt(rho) (TightBinding.cxx:133-143)
__device__ __forceinline__ void TBModel::embed(const Table* T, double rho, double& F, double& Fp) {
  F = -T->c * sqrt(rho);
  Fp = (rho != 0.0) ? -T->c / (2.0 * sqrt(rho)) : 0.0;
}

template <class M>
__global__ void __launch_bounds__(256) k_eam_density(Atoms A, Box bx, const typename M::Table* __restrict__ T,
                                                     const signed char* __restrict__ ek, double2* __restrict__ fe,
                                                     double* __restrict__ eatom, int* __restrict__ row,
                                                     int* __restrict__ rcnt, unsigned* err) {
  constexpr int LW = 16;
  const int lane = threadIdx.x & 63;
  const int sl = lane % LW;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mine = ((1ull << LW) - 1ull) << (lane / LW * LW);
  const int i = blockIdx.x * (blockDim.x / LW) + threadIdx.x / LW;
  if (i >= A.n) return;  // group-uniform
  const int ei = ek[i];
  if (ei < 0) {
    if (sl == 0) {
      fe[i] = make_double2(0.0, -1.0);
      rcnt[i] = 0;
    }
    return;
  }
  int cnt = A.ml_cnt[i];  // the metal row, or the long row when it did not fit
  const int* L = A.ml + (size_t)i * kMetalRow;
  if (cnt < 0) {
    cnt = A.nll_cnt[i];
    L = A.nll + (size_t)i * A.cap_l;
  }
  int* my = row + (size_t)i * kEamRow;
  const double4 pi = A.p[i];
  const Frame fr = frame_of(bx, i);
  const double rmax = M::rmax(T);
  double rho = 0.0;
  int m = 0;
  for (int b0 = 0; b0 < cnt; b0 += LW) {  // group-uniform trip count
    const int q = b0 + sl;
    bool in = false;
    int j = 0;
    if (q < cnt) {
      j = L[q];
      const double4 pj = A.p[j];
      if ((int)pj.w == T_OTHER) {  // the metals are the atoms of type T_OTHER
        const D3 dd = disp(bx, fr, i, j, pi, pj);
        const double dx = dd.x, dy = dd.y, dz = dd.z;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r < rmax) {
          const int ej = ek[j];
          if (ej >= 0) {
            in = true;
            double g, dg;
            M::g(T, ei, ej, r, g, dg);
            rho += g;
          }
        }
      }
    }
    // the in-range partners, in list order, for the force pass
    const unsigned long long msk = __ballot(in) & mine;
    if (in) {
      const int slot = m + __popcll(msk & below);
      if (slot < kEamRow) my[slot] = j;
    }
    m += __popcll(msk);
  }
#pragma unroll
  for (int o = LW / 2; o > 0; o >>= 1) rho += __shfl_xor(rho, o, LW);
  if (sl == 0) {
    double F, Fp;
    M::embed(T, rho, F, Fp);
    fe[i] = make_double2(Fp, double(ei));
    eatom[i] = F;
    if (m > kEamRow) raise(sys_err(err, A.sys[i]), ERR_EAM_ROW_OVERFLOW);
    rcnt[i] = m < kEamRow ? m : kEamRow;
  }
}

// Each pair is visited from both atoms (gather), so each visit adds half of
// the pair's virial.
template <class M, bool VIR>
__global__ void __launch_bounds__(256) k_eam_force(Atoms A, Box bx, const typename M::Table* __restrict__ T,
                                                   const signed char* __restrict__ ek,
                                                   const double2* __restrict__ fe, const int* __restrict__ row,
                                                   const int* __restrict__ rcnt, double* __restrict__ grad,
                                                   double* __restrict__ eatom, long long* __restrict__ vir) {
  constexpr int LW = 16;
  const int sl = threadIdx.x % LW;
  const int i = blockIdx.x * (blockDim.x / LW) + threadIdx.x / LW;
  if (i >= A.n) return;  // group-uniform
  const int ei = ek[i];
  if (ei < 0) return;
  const int cnt = rcnt[i];
  const int* L = row + (size_t)i * kEamRow;
  const double4 pi = A.p[i];
  const Frame fr = frame_of(bx, i);
  const double fpi = fe[i].x;
  double gx = 0.0, gy = 0.0, gz = 0.0, ev = 0.0;
  Virial<VIR> vw(vir, 1.0);
  if constexpr (VIR) vw.at(A.sys[i]);
  for (int q = sl; q < cnt; q += LW) {
    const int j = L[q];
    const double4 pj = A.p[j];
    const double2 fj = fe[j];  // F'(rho_j) and j's element
    const D3 dd = disp(bx, fr, i, j, pi, pj);
        const double dx = dd.x, dy = dd.y, dz = dd.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const int ej = (int)fj.y;
    double v, dv, g, dg;
    M::pair(T, ei, ej, r, v, dv);
    M::g(T, ei, ej, r, g, dg);
    const double c = (dv + (fpi + fj.x) * dg) / r;
    gx += c * dx;
    gy += c * dy;
    gz += c * dz;
    vw.push_half(D3{dx, dy, dz}, D3{c * dx, c * dy, c * dz});
    if (j > i) ev += v;
  }
  vw.template flush_group<LW>(sl);
#pragma unroll
  for (int o = LW / 2; o > 0; o >>= 1) {
    gx += __shfl_xor(gx, o, LW);
    gy += __shfl_xor(gy, o, LW);
    gz += __shfl_xor(gz, o, LW);
    ev += __shfl_xor(ev, o, LW);
  }
  if (sl == 0) {
    atomic_add3(GradSink{grad, A.gfix}, i, D3{gx, gy, gz});
    eatom[i] += ev;
  }
}

template __global__ void k_eam_density<AcklandModel>(Atoms, Box, const EAMTable*, const signed char*, double2*,
                                                     double*, int*, int*, unsigned*);
template __global__ void k_eam_force<AcklandModel, false>(Atoms, Box, const EAMTable*, const signed char*,
                                                          const double2*, const int*, const int*, double*, double*,
                                                          long long*);
template __global__ void k_eam_force<AcklandModel, true>(Atoms, Box, const EAMTable*, const signed char*,
                                                         const double2*, const int*, const int*, double*, double*,
                                                         long long*);
template __global__ void k_eam_density<TBModel>(Atoms, Box, const TBTable*, const signed char*, double2*, double*,
                                                int*, int*, unsigned*);
template __global__ void k_eam_force<TBModel, false>(Atoms, Box, const TBTable*, const signed char*,
                                                     const double2*, const int*, const int*, double*, double*,
                                                     long long*);
template __global__ void k_eam_force<TBModel, true>(Atoms, Box, const TBTable*, const signed char*,
                                                    const double2*, const int*, const int*, double*, double*,
                                                    long long*);

// FBM::operator() (FBM.cxx:70-94) for the pairs Cu-Cu and Cu-Ar: 16 lanes
// per Cu or Ar atom gather its handled partners from the long list (both
// directions, no scatter to j); the pair's energy goes to its lower-index
// atom, as the reference's i < j loop sums it.
__device__ __forceinline__ int fbm_index(int elem) { return elem == 64 ? 0 : elem == 40 ? 1 : -1; }

template <bool VIR>
__global__ void __launch_bounds__(256) k_fbm(Atoms A, Box bx, const FBMTable* __restrict__ T,
                                             double* __restrict__ grad, double* __restrict__ eatom,
                                             long long* __restrict__ vir) {
  constexpr int LW = 16;
  const int sl = threadIdx.x % LW;
  const int i = blockIdx.x * (blockDim.x / LW) + threadIdx.x / LW;
  if (i >= A.n) return;  // group-uniform
  const int ei = fbm_index(A.elem[i]);
  if (ei < 0) return;
  const int cnt = A.nll_cnt[i];
  const int* L = A.nll + (size_t)i * A.cap_l;
  const double4 pi = A.p[i];
  const Frame fr = frame_of(bx, i);
  const double R0 = T->R0, R1 = T->R1;
  double gx = 0.0, gy = 0.0, gz = 0.0, ev = 0.0;
  Virial<VIR> vw(vir, 1.0);
  if constexpr (VIR) vw.at(A.sys[i]);
  for (int q = sl; q < cnt; q += LW) {
    const int j = L[q];
    const double4 pj = A.p[j];
    const D3 dd = disp(bx, fr, i, j, pi, pj);
        const double dx = dd.x, dy = dd.y, dz = dd.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > R1 * R1) continue;  // FPairwise::probablyAreNeighbours
    const int ej = fbm_index(A.elem[j]);
    if (ej < 0 || (ei == 1 && ej == 1)) continue;  // Cu-Cu and Cu-Ar only
    const double r = sqrt(d2);
    double f, df;
    cutoff(r, R0, R1, f, df);
    const double e = T->A[ei][ej] * exp(-T->B[ei][ej] * r);
    const double c = (-T->B[ei][ej] * e * f + e * df) / r;
    gx += c * dx;
    gy += c * dy;
    gz += c * dz;
    vw.push_half(D3{dx, dy, dz}, D3{c * dx, c * dy, c * dz});
    if (j > i) ev += e * f;
  }
  vw.template flush_group<LW>(sl);
#pragma unroll
  for (int o = LW / 2; o > 0; o >>= 1) {
    gx += __shfl_xor(gx, o, LW);
    gy += __shfl_xor(gy, o, LW);
    gz += __shfl_xor(gz, o, LW);
    ev += __shfl_xor(ev, o, LW);
  }
  if (sl == 0) {
    atomic_add3(GradSink{grad, A.gfix}, i, D3{gx, gy, gz});
    eatom[i] = ev;
  }
}
template __global__ void k_fbm<false>(Atoms, Box, const FBMTable*, double*, double*, long long*);
template __global__ void k_fbm<true>(Atoms, Box, const FBMTable*, double*, double*, long long*);

// ------------------------------------------------------------ Verlet
// eq.1 of SimLoop::executeMain (SimLoop.cxx:323-332): x += v dt + a dt^2/2,
// accumulated into both neighbour lists' (and the LJ pair rows') displacements;
// the lengths of the atom's accumulated displacements come back for the skin
// test.
__device__ __forceinline__ void drift_atom(const Atoms& A, int i, double dt, bool live, double& ls, double& ll,
                                           double& lp, double& le) {
  double sx = A.disp_s[3 * i], sy = A.disp_s[3 * i + 1], sz = A.disp_s[3 * i + 2];
  double lx = A.disp_l[3 * i], ly = A.disp_l[3 * i + 1], lz = A.disp_l[3 * i + 2];
  double px = 0, py = 0, pz = 0, ex = 0, ey = 0, ez = 0;
  if (A.disp_p) {
    px = A.disp_p[3 * i];
    py = A.disp_p[3 * i + 1];
    pz = A.disp_p[3 * i + 2];
  }
  if (A.disp_e) {
    ex = A.disp_e[3 * i];
    ey = A.disp_e[3 * i + 1];
    ez = A.disp_e[3 * i + 2];
  }
  if (live && !A.fixed[i]) {
    double dx = A.vx[i] * dt + A.ax[i] * dt * dt / 2.0;
    double dy = A.vy[i] * dt + A.ay[i] * dt * dt / 2.0;
    double dz = A.vz[i] * dt + A.az[i] * dt * dt / 2.0;
    double x = A.x[i] + dx, y = A.y[i] + dy, z = A.z[i] + dz;
    A.x[i] = x;
    A.y[i] = y;
    A.z[i] = z;
    A.p[i] = packed_pos(A, i, x, y, z);
    sx += dx;
    sy += dy;
    sz += dz;
    lx += dx;
    ly += dy;
    lz += dz;
    A.disp_s[3 * i] = sx;
    A.disp_s[3 * i + 1] = sy;
    A.disp_s[3 * i + 2] = sz;
    A.disp_l[3 * i] = lx;
    A.disp_l[3 * i + 1] = ly;
    A.disp_l[3 * i + 2] = lz;
    if (A.disp_p) {
      px += dx;
      py += dy;
      pz += dz;
      A.disp_p[3 * i] = px;
      A.disp_p[3 * i + 1] = py;
      A.disp_p[3 * i + 2] = pz;
    }
    if (A.disp_e) {
      ex += dx;
      ey += dy;
      ez += dz;
      A.disp_e[3 * i] = ex;
      A.disp_e[3 * i + 1] = ey;
      A.disp_e[3 * i + 2] = ez;
    }
  }
  ls = sqrt(sx * sx + sy * sy + sz * sz);
  ll = sqrt(lx * lx + ly * ly + lz * lz);
  lp = sqrt(px * px + py * py + pz * pz);
  le = sqrt(ex * ex + ey * ey + ez * ez);
}

// The drift of one tile's atoms and the tile's part of the skin test (the
// two largest displacement lengths per list), as k_sys_disp_part writes it.
// A system whose simulated time has reached its final time stands still
// (the loop condition of SimLoop.cxx:252).
__global__ void __launch_bounds__(256) k_drift_tile(Atoms A, const int2* __restrict__ rb,
                                                    const double* __restrict__ dtp, const double* __restrict__ simt,
                                                    const double* __restrict__ tfin, double* __restrict__ part) {
  const int2 t = rb[blockIdx.x];
  const int e = tile_end(t, A.sys_off);
  const double dt = dtp[t.x];
  const bool live = simt[t.x] < tfin[t.x];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = t.y + threadIdx.x; i < e; i += blockDim.x) {
    double ls, ll, lp, le;
    drift_atom(A, i, dt, live, ls, ll, lp, le);
    top2_insert(ls, v[0], v[1]);
    top2_insert(ll, v[2], v[3]);
    if (A.disp_p) top2_insert(lp, v[4], v[5]);
    if (A.disp_e) top2_insert(le, v[6], v[7]);
  }
  top2_tile_out(v, part);
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

// Membership of the bath geometry (SimLoop.hpp:139-151,164-170,174-185) at
// the given (unwrapped) position; L is the system's box.
__device__ __forceinline__ bool bath_inside(const Bath& b, double x, double y, double z, const double* L) {
  switch (b.geom) {
    case TB_UNIVERSE:
      return true;
    case TB_BOX:
      return (z > b.zMin) ||
             (((x < 0.0 + b.dBoundary) || (x > L[0] - b.dBoundary) || (y < 0.0 + b.dBoundary) ||
               (y > L[1] - b.dBoundary)) &&
              z > b.zFree);
    case TB_SPHERE: {
      double dx = b.cx - x, dy = b.cy - y, dz = b.cz - z;
      return z > b.zFree && sqrt(dx * dx + dy * dy + dz * dz) > b.radius;
    }
    default:
      return false;
  }
}

// actualTemperatureOfThermalBath (SimLoop.cxx:516-535), tiled per system, at
// the start of the step (positions and velocities at t).
__global__ void __launch_bounds__(256) k_bath_temp_part(Atoms A, Box bx, const int2* __restrict__ rb,
                                                        double* __restrict__ part) {
  const int2 t = rb[blockIdx.x];
  const int s = t.x;
  const Bath b = A.bath[s];
  double e = 0.0, cnt = 0.0;
  if (b.geom != TB_NONE) {
    for (int i = t.y + threadIdx.x; i < tile_end(t, A.sys_off); i += blockDim.x) {
      if (A.fixed[i] || !A.tb_apply[i] || !bath_inside(b, A.x[i], A.y[i], A.z[i], own_pbc(bx, i).v)) continue;
      double vv = sqrt(A.vx[i] * A.vx[i] + A.vy[i] * A.vy[i] + A.vz[i] * A.vz[i]);
      e += A.mass[i] * (vv * vv) / 2.0;
      cnt += 1.0;
    }
  }
  e = block_sum(e);
  cnt = block_sum(cnt);
  if (threadIdx.x == 0) {
    part[8 * (size_t)blockIdx.x] = e;
    part[8 * (size_t)blockIdx.x + 1] = cnt;
  }
}

__global__ void __launch_bounds__(64) k_bath_temp_merge(Atoms A, const int* __restrict__ rb_off,
                                                        const double* __restrict__ part) {
  const int s = blockIdx.x;
  double e = 0.0, cnt = 0.0;
  for (int q = rb_off[s] + threadIdx.x; q < rb_off[s + 1]; q += 64) {
    e += part[8 * (size_t)q];
    cnt += part[8 * (size_t)q + 1];
  }
  e = wave_sum(e);
  cnt = wave_sum(cnt);
  if (threadIdx.x == 0)
    A.tb_T[s] = (A.bath[s].geom == TB_NONE || cnt == 0.0) ? 0.0 : e / (3.0 / 2.0 * (8.617343E-5 * kEV) * cnt);
}

// eq.2-4 of SimLoop::executeMain (SimLoop.cxx:338-399) with the thermal
// bath's friction (:352-386) on atoms inside its geometry at x(t+dt), then
// the PBC wrap (:401, Atom.cxx:93-125).
// One atom: v (in: v(t), out: v(t+dt)) from grad g; dE = the energy taken
// from the bath (0 outside it).
__device__ __forceinline__ void kick_atom(const Atoms& A, const Box& bx, int i, double dt, double gx, double gy,
                                          double gz, double& vx, double& vy, double& vz, double& dE) {
  const Pbc3 own = own_pbc(bx, i);  // the atom's own PBC (Atom::applyPBC wraps by it)
  const double* L = own.v;
  double x = A.x[i], y = A.y[i], z = A.z[i];
  dE = 0.0;
  if (!A.fixed[i]) {
    const int s = A.sys[i];
    double M = A.mass[i];
    double fx = -gx, fy = -gy, fz = -gz;
    double hx = vx + A.ax[i] * dt / 2.0;
    double hy = vy + A.ay[i] * dt / 2.0;
    double hz = vz + A.az[i] * dt / 2.0;
    if (A.bath) {
      const Bath& b = A.bath[s];
      if (b.geom != TB_NONE && A.tb_apply[i] && bath_inside(b, x, y, z, L)) {
        const double T = A.tb_T[s];
        double To_by_T = (fabs(T) < 1e-5) ? 0 : (b.To / T);
        if (To_by_T < -5.0) To_by_T = -5.0;
        if (To_by_T > 5.0) To_by_T = 5.0;
        const double k = 1.0 - sqrt(To_by_T);
        const double dfx = -vx * M * b.gamma * k, dfy = -vy * M * b.gamma * k, dfz = -vz * M * b.gamma * k;
        // energy taken from the bath: kinetic energy after the step with and
        // without its force (eq.2 and eq.4 with both accelerations)
        const double nx = fx / M, ny = fy / M, nz = fz / M;
        const double ux = A.an_nt[3 * i] * dt / 2.0 + nx * dt / 2.0;
        const double uy = A.an_nt[3 * i + 1] * dt / 2.0 + ny * dt / 2.0;
        const double uz = A.an_nt[3 * i + 2] * dt / 2.0 + nz * dt / 2.0;
        A.an_nt[3 * i] = nx;
        A.an_nt[3 * i + 1] = ny;
        A.an_nt[3 * i + 2] = nz;
        const double bx_ = (fx + dfx) / M, by_ = (fy + dfy) / M, bz_ = (fz + dfz) / M;
        const double wx = A.ax[i] * dt / 2.0 + bx_ * dt / 2.0;
        const double wy = A.ay[i] * dt / 2.0 + by_ * dt / 2.0;
        const double wz = A.az[i] * dt / 2.0 + bz_ * dt / 2.0;
        const double r1 = sqrt((vx + wx) * (vx + wx) + (vy + wy) * (vy + wy) + (vz + wz) * (vz + wz));
        const double r0 = sqrt((vx + ux) * (vx + ux) + (vy + uy) * (vy + uy) + (vz + uz) * (vz + uz));
        dE = M * (r1 * r1) / 2.0 - M * (r0 * r0) / 2.0;
        fx += dfx;
        fy += dfy;
        fz += dfz;
      }
      A.tb_dE[i] = dE;
    }
    double ax = fx / M, ay = fy / M, az = fz / M;
    A.ax[i] = ax;
    A.ay[i] = ay;
    A.az[i] = az;
    vx = hx + ax * dt / 2.0;
    vy = hy + ay * dt / 2.0;
    vz = hz + az * dt / 2.0;
    A.vx[i] = vx;
    A.vy[i] = vy;
    A.vz[i] = vz;
  }
  int cx = A.pcx[i], cy = A.pcy[i], cz = A.pcz[i];
  wrap(x, L[0], cx);
  wrap(y, L[1], cy);
  wrap(z, L[2], cz);
  A.x[i] = x;
  A.y[i] = y;
  A.z[i] = z;
  A.pcx[i] = cx;
  A.pcy[i] = cy;
  A.pcz[i] = cz;
  A.p[i] = packed_pos(A, i, x, y, z);
}

// The kick of one tile's atoms and the tile's partial step outputs: the
// potential energy (all chain members), the kinetic energy after the kick,
// the net force, the bath's energy, and the largest speed of a free atom with
// neighbours (the dt rule); merged per system by k_sys_step_merge.
__global__ void __launch_bounds__(256) k_kick_tile(Atoms A, Box bx, const int2* __restrict__ rb,
                                                   const double* __restrict__ dtp, const double* __restrict__ simt,
                                                   const double* __restrict__ tfin, int nterms,
                                                   const double* __restrict__ eatom, double* __restrict__ part) {
  const int2 t = rb[blockIdx.x];
  const int e = tile_end(t, A.sys_off);
  const double dt = dtp[t.x];
  const bool live = simt[t.x] < tfin[t.x];
  const int n = A.n;
  double ep = 0, ek = 0, fx = 0, fy = 0, fz = 0, vm = 0, eb = 0;
  for (int i = t.y + threadIdx.x; i < e; i += blockDim.x) {
    const double gx = A.grad[3 * i], gy = A.grad[3 * i + 1], gz = A.grad[3 * i + 2];
    double vx = A.vx[i], vy = A.vy[i], vz = A.vz[i], dE = 0.0;
    if (live) kick_atom(A, bx, i, dt, gx, gy, gz, vx, vy, vz, dE);
    for (int q = 0; q < nterms; q++) ep += eatom[(size_t)q * n + i];
    double vv = sqrt(vx * vx + vy * vy + vz * vz);
    ek += A.mass[i] * (vv * vv) / 2.0;
    fx -= gx;
    fy -= gy;
    fz -= gz;
    // FProxy::hasNeighbors (SimLoop.cxx:394-398): the short list's members
    // handle H and C only, the long-range members by their own element sets
    const bool nb = A.nls_cnt[i] > 0 || (A.nbl ? A.nbl[i] != 0 : A.nll_cnt[i] > 0);
    if (!A.fixed[i] && nb && vv > vm) vm = vv;
    eb += dE;  // energy taken from the bath this step (SimLoop.cxx:384)
  }
  ep = block_sum(ep);
  ek = block_sum(ek);
  fx = block_sum(fx);
  fy = block_sum(fy);
  fz = block_sum(fz);
  eb = block_sum(eb);
  for (int o = 32; o > 0; o >>= 1) vm = fmax(vm, __shfl_xor(vm, o, 64));
  __shared__ double vms[4];
  if ((threadIdx.x & 63) == 0) vms[threadIdx.x >> 6] = vm;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) vm = fmax(vm, vms[q]);
    double* o = part + 8 * (size_t)blockIdx.x;
    o[0] = ep; o[1] = ek; o[2] = fx; o[3] = fy; o[4] = fz; o[5] = vm; o[6] = eb;
  }
}

// Refreshes the packed positions after the host wrote x, y, z.
__global__ void k_pack_pos(Atoms A) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A.n) A.p[i] = packed_pos(A, i, A.x[i], A.y[i], A.z[i]);
}

// First force evaluation (iteration 0, SimLoop.cxx:305-321): a = -grad/M,
// except in the systems whose accelerations were loaded (keep[s] = 1).
__global__ void k_init_acc(Atoms A, const unsigned char* __restrict__ keep) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n || A.fixed[i]) return;
  if (keep && keep[A.sys[i]]) return;
  double M = A.mass[i];
  A.ax[i] = -A.grad[3 * i] / M;
  A.ay[i] = -A.grad[3 * i + 1] / M;
  A.az[i] = -A.grad[3 * i + 2] / M;
  A.an_nt[3 * i] = A.ax[i];
  A.an_nt[3 * i + 1] = A.ay[i];
  A.an_nt[3 * i + 2] = A.az[i];
}

// Per-system energies of each chain member: eacc[8*s+t] = sum of the
// per-atom energies eatom[t*n+i] over the system's atoms.
__global__ void k_sys_energy(int nterms, const int* __restrict__ sys_off, int n, const double* __restrict__ eatom,
                             double* __restrict__ eacc) {
  int s = blockIdx.x;
  int b = sys_off[s], e = sys_off[s + 1];
  for (int t = 0; t < nterms; t++) {
    double v = 0;
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) v += eatom[(size_t)t * n + i];
    v = block_sum(v);
    if (threadIdx.x == 0) eacc[8 * s + t] = v;
  }
}

// Per-system end of an iteration: E_pot (t+dt), E_kin after the kick
// (SimLoop.cxx:500-514), |sum F| (:418-430), the dt rule over atoms with
// neighbours (:394-398, :409-416), tiled per system: the tile pass is
// k_kick_tile.
// One wave per system: out[(step*nsys + s)*4 + {0,1,2,3}] = epot, ekin, dt
// used (0 once the system has reached its final time), |sum F|; the bath's
// energy; the simulated time and the next dt.
__global__ void __launch_bounds__(64) k_sys_step_merge(const int* __restrict__ rb_off, const double* __restrict__ part,
                                                       double* __restrict__ dtp, int fixed_dt, double* __restrict__ out,
                                                       const int* __restrict__ stepidx, int nsys,
                                                       double* __restrict__ ebath, double* __restrict__ simt,
                                                       double* __restrict__ tfin, double* __restrict__ dtprev,
                                                       long long* __restrict__ iter, const unsigned* __restrict__ serr) {
  const int s = blockIdx.x;
  double ep = 0, ek = 0, fx = 0, fy = 0, fz = 0, vm = 0, eb = 0;
  for (int q = rb_off[s] + threadIdx.x; q < rb_off[s + 1]; q += 64) {
    const double* o = part + 8 * (size_t)q;
    ep += o[0]; ek += o[1]; fx += o[2]; fy += o[3]; fz += o[4]; vm = fmax(vm, o[5]); eb += o[6];
  }
  ep = wave_sum(ep);
  ek = wave_sum(ek);
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  eb = wave_sum(eb);
  for (int o = 32; o > 0; o >>= 1) vm = fmax(vm, __shfl_xor(vm, o, 64));
  if (threadIdx.x == 0) {
    if (ebath) ebath[s] += eb;
    int step = *stepidx;
    double* o = out + ((size_t)step * nsys + s) * 4;
    o[0] = ep;
    o[1] = ek;
    const bool live = simt[s] < tfin[s];  // a system past its final time stood still this step
    o[2] = live ? dtp[s] : 0.0;
    o[3] = sqrt(fx * fx + fy * fy + fz * fz);
    if (live) {
      simt[s] += dtp[s];  // SimLoop.cxx:405
      dtprev[s] = dtp[s];  // :407
      iter[s] += 1;
      if (!fixed_dt) {
        const double dt_max = 5e-16, dt_min = 1e-20;
        double dt;
        if (vm != 0.0) dt = 0.05 * (1.0 * kAo) / vm;
        else dt = dt_max;
        if (dt > dt_max) dt = dt_max;
        if (dt < dt_min) dt = dt_min;
        dtp[s] = dt;
      }
    }
    // a device error in this system's terms (a capacity or a range check)
    // stops it here: it stands still from the next step on, the others go on
    if (serr[s]) tfin[s] = simt[s];
  }
}

// A block per system, before the iteration's drift (doBeforeIteration and
// the XVA check run first in the loop body, SimLoop.cxx:252-277).  Due
// tests on the clock at the iteration's start, t, and the previous step's
// dt: t == 0 or int(t / I) != int((t - dt_prev) / I), with I the SnapshotList
// interval of CustomSimLoop::doBeforeIteration (mdtrajsim/main.cpp:58-83: 1e-14
// s below 2 ps, 5e-14 s below 4 ps, 1e-13 s beyond) or the XVA interval it
// sets (1 ps below 2 ps, 5 ps beyond; iteration 0 always writes one,
// SimLoop.cxx:271-277).
__global__ void __launch_bounds__(256) k_traj_capture(Atoms A, TrajBuf tb, const double* __restrict__ simt,
                                                      const double* __restrict__ tfin,
                                                      const double* __restrict__ dtprev,
                                                      const long long* __restrict__ iter,
                                                      const double* __restrict__ dtp) {
  __shared__ long long s_first[2];
  __shared__ int s_n[2];
  const int s = blockIdx.x;
  const double t = simt[s];
  if (!(t < tfin[s])) return;  // block-uniform: the system has stopped
  if (threadIdx.x == 0) {
    const double dp = dtprev[s];
    auto due = [&](double I) { return t == 0.0 || (long long)(t / I) != (long long)((t - dp) / I); };
    const double isnap = t < 2e-12 ? 5e-16 * 5 * 4 : t < 4e-12 ? 5e-16 * 25 * 4 : 5e-16 * 50 * 4;
    const double ixva = t < 2e-12 ? 1e-12 : 5e-12;
    const int nsel = tb.sel_off ? tb.sel_off[s + 1] - tb.sel_off[s] : 0;
    const int ns = A.sys_off[s + 1] - A.sys_off[s];
    const int want[2] = {(nsel > 0 && due(isnap)) ? nsel : 0, (tb.xva && (iter[s] == 0 || due(ixva))) ? ns : 0};
    for (int k = 0; k < 2; k++) {
      s_n[k] = 0;
      if (!want[k]) continue;
      const unsigned long long f = atomicAdd(&tb.cnt[0], 1ull);
      const unsigned long long r = atomicAdd(&tb.cnt[1], (unsigned long long)want[k]);
      if ((long long)f >= tb.cap_frm || (long long)(r + want[k]) > tb.cap_rec) {
        atomicMax(&tb.cnt[2], 1ull);
        continue;
      }
      TrajFrame& F = tb.frm[f];
      F.kind = k == 0 ? TRAJ_SNAPSHOT : TRAJ_XVA;
      F.system = s;
      F.iteration = iter[s];
      F.sim_time = t;
      F.dt = dtp[s];
      F.first = (long long)r;
      F.count = want[k];
      s_n[k] = want[k];
      s_first[k] = (long long)r;
    }
  }
  __syncthreads();
  for (int k = 0; k < 2; k++)
    for (int q = threadIdx.x; q < s_n[k]; q += blockDim.x) {
      const int i = (k == 0) ? tb.sel[tb.sel_off[s] + q] : A.sys_off[s] + q;
      TrajRec& R = tb.rec[s_first[k] + q];
      R.x[0] = A.x[i];
      R.x[1] = A.y[i];
      R.x[2] = A.z[i];
      R.v[0] = A.vx[i];
      R.v[1] = A.vy[i];
      R.v[2] = A.vz[i];
      R.pc[0] = A.pcx[i];
      R.pc[1] = A.pcy[i];
      R.pc[2] = A.pcz[i];
      R.atom = i - A.sys_off[s];
    }
}

// hasNeighbors for the long-range members after a long-list build: 16 lanes
// per atom scan its row (positions as the build saw them) for a partner that
// some member handling both elements lists.
__global__ void __launch_bounds__(256) k_long_nbflag(const int* __restrict__ flag, const int* __restrict__ sflags,
                                                     Atoms A, Box bx, LongMembers lm) {
  if (!*flag) return;  // the long list's rebuild flag
  constexpr int LW = 16;
  const int sl = threadIdx.x % LW;
  const int i = blockIdx.x * (blockDim.x / LW) + threadIdx.x / LW;
  if (i >= A.n) return;  // group-uniform
  if (!sflags[2 * A.sys[i] + 1]) return;
  const unsigned ei = elem_bit(A.elem[i]);
  unsigned mine = 0;  // members that handle atom i
  for (int q = 0; q < lm.m; q++)
    if (lm.handled[q] & ei) mine |= 1u << q;
  bool found = false;
  if (mine) {
    const int cnt = A.nll_cnt[i];
    const int* L = A.nll + (size_t)i * A.cap_l;
    const double4 pi = A.p[i];
    const Frame fr = frame_of(bx, i);
    for (int b0 = 0; b0 < cnt && !found; b0 += LW) {
      const int q = b0 + sl;
      bool hit = false;
      if (q < cnt) {
        const int j = L[q];
        const unsigned ej = elem_bit(A.elem[j]);
        const double4 pj = A.p[j];
        const D3 dd = disp(bx, fr, i, j, pi, pj);
        const double dx = dd.x, dy = dd.y, dz = dd.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        for (int m = 0; m < lm.m; m++)
          if (((mine >> m) & 1u) && (lm.handled[m] & ej) && d2 < lm.r2[m]) hit = true;
      }
      const unsigned long long grp = ((1ull << LW) - 1ull) << ((threadIdx.x & 63) / LW * LW);
      found = (__ballot(hit) & grp) != 0ull;
    }
  }
  if (sl == 0) A.nbl[i] = found ? 1 : 0;
}

// The end of a force call of the FP64 form: the fixed-point gradient into
// grad and the fixed-point energy slots (bit t of fixmask) into eatom.
__global__ void k_fix_join(Atoms A, int nterms, unsigned fixmask, double* __restrict__ eatom) {
  const size_t n = (size_t)A.n;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < 3 * n + nterms * n;
       q += (size_t)gridDim.x * blockDim.x) {
    if (q < 3 * n) {
      A.grad[q] = fix_join(A.gfix[2 * q], A.gfix[2 * q + 1]);
    } else {
      const size_t e = q - 3 * n;
      if ((fixmask >> (e / n)) & 1u) eatom[e] = fix_join(A.efix[2 * e], A.efix[2 * e + 1]) * (1.0 / kFixEnergy);
    }
  }
}

// The force call's per-system virial into this step's output slot.
__global__ void k_step_virial(int nsys, const double* __restrict__ vir, double* __restrict__ stepvir,
                              const int* __restrict__ stepidx) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < 6 * nsys) stepvir[(size_t)(*stepidx) * 6 * nsys + q] = vir[q];
}

// The virial of a force call from its fixed-point sums (erg).
__global__ void k_vir_join(int nsys, const long long* __restrict__ fix, double* __restrict__ vir) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < 6 * nsys) vir[q] = fix_join(fix[2 * q], fix[2 * q + 1]) * (1.0 / kFixEnergy);
}

__global__ void k_step_advance(int* __restrict__ stepidx) {
  if (threadIdx.x == 0) *stepidx += 1;
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
