// Neighbour lists, skin test, ZBL and the velocity-Verlet kernels (FP64).
// Kernels are compiled without FMA contraction so that a*b+c rounds twice, as
// in the reference's x86-64 build.
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

// Merge of two top-2 sets (the two largest values of the union are exact
// whatever the merge order).
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

__global__ void k_disp_top2(int n, const double* __restrict__ ds, const double* __restrict__ dl,
                            double* __restrict__ part) {
  __shared__ double sh[4][4];
  double a1 = 0, a2 = 0, b1 = 0, b2 = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double m = sqrt(ds[3 * i] * ds[3 * i] + ds[3 * i + 1] * ds[3 * i + 1] + ds[3 * i + 2] * ds[3 * i + 2]);
    top2_insert(m, a1, a2);
    double q = sqrt(dl[3 * i] * dl[3 * i] + dl[3 * i + 1] * dl[3 * i + 1] + dl[3 * i + 2] * dl[3 * i + 2]);
    top2_insert(q, b1, b2);
  }
  top2_wave(a1, a2);
  top2_wave(b1, b2);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w][0] = a1; sh[w][1] = a2; sh[w][2] = b1; sh[w][3] = b2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      top2_merge(x1, x2, sh[q][0], sh[q][1]);
      top2_merge(y1, y2, sh[q][2], sh[q][3]);
    }
    part[4 * blockIdx.x + 0] = x1;
    part[4 * blockIdx.x + 1] = x2;
    part[4 * blockIdx.x + 2] = y1;
    part[4 * blockIdx.x + 3] = y2;
  }
}

// One 256-thread block merges the per-block top-2 values and sets the flags
// (rebuild iff d1 + d2 > 0.44 Rc).
__global__ void k_disp_decide(int nblk, const double* __restrict__ part, double lim_s, double lim_l,
                              int* __restrict__ flags, long long* __restrict__ rebuilds) {
  __shared__ double sh[4][4];
  double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
  for (int q = threadIdx.x; q < nblk; q += blockDim.x) {
    top2_merge(x1, x2, part[4 * q], part[4 * q + 1]);
    top2_merge(y1, y2, part[4 * q + 2], part[4 * q + 3]);
  }
  top2_wave(x1, x2);
  top2_wave(y1, y2);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w][0] = x1; sh[w][1] = x2; sh[w][2] = y1; sh[w][3] = y2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a1 = 0, a2 = 0, b1 = 0, b2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      top2_merge(a1, a2, sh[q][0], sh[q][1]);
      top2_merge(b1, b2, sh[q][2], sh[q][3]);
    }
    flags[0] |= (a1 + a2 > lim_s) ? 1 : 0;
    flags[1] |= (b1 + b2 > lim_l) ? 1 : 0;
    if (flags[0]) rebuilds[0] += 1;
    if (flags[1]) rebuilds[1] += 1;
  }
}

__global__ void k_cell_clear(const int* __restrict__ flag, int ncell, int* __restrict__ cnt) {
  if (!*flag) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += gridDim.x * blockDim.x) cnt[i] = 0;
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
                           int fixed_dt, double* __restrict__ out, int* __restrict__ stepidx) {
  if (threadIdx.x != 0) return;
  int step = *stepidx;
  *stepidx = step + 1;
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
