// Brenner (1990) bond-order potential, FP64, one bond per thread with a
// forward pass and an adjoint pass (Brenner::operator(), Brenner.cxx:31-61;
// Baver :129-141, D :168-196, B :198-209, G :211-243, Nt/NH/NC :245-298,
// Nconj :352-410, HN :412-442, FN :444-465).  Unlike REBO, the coordination
// numbers are recomputed as sums over the neighbour list, in list order.
#include "mdg_kernels.hpp"

namespace mdg {

namespace {
struct BRef {
  int k;
  double r, f, df;
  D3 rv;
};
__device__ __forceinline__ BRef bref(const Atoms& A, int a, int s) {
  int o = a * A.maxb + s;
  BondRec<double> q = A.b64[o];
  BRef b;
  b.k = A.bj[o];
  b.r = q.r;
  b.f = q.f;
  b.df = q.df;
  b.rv = d3(q.x, q.y, q.z);
  return b;
}

// Sum of f_ak over a's neighbours other than b, optionally restricted to one
// type (Brenner::Nt / NH / NC).
__device__ __forceinline__ double nsum(const Atoms& A, int a, int b, int only_type) {
  double s = 0;
  int nb = A.bcnt[a];
  for (int t = 0; t < nb; t++) {
    BRef k = bref(A, a, t);
    if (k.k == b) continue;
    if (only_type >= 0 && A.type[k.k] != only_type) continue;
    s += k.f;
  }
  return s;
}

// Scatter c * d f_ak/dx over a's neighbours other than b (optionally one type).
template <class Wv>
__device__ __forceinline__ void nsum_back(const Atoms& A, int a, int b, int only_type, double c, D3& ga,
                                          double* grad, Wv& vw) {
  if (c == 0.0) return;
  int nb = A.bcnt[a];
  for (int t = 0; t < nb; t++) {
    BRef k = bref(A, a, t);
    if (k.k == b || k.df == 0.0) continue;
    if (only_type >= 0 && A.type[k.k] != only_type) continue;
    D3 v = k.rv * (k.df * c / k.r);
    ga = ga + v;
    vw.push(k.rv, v);
    atomic_add3(GradSink{grad, A.gfix}, k.k, -v);
  }
}

__device__ __forceinline__ double Fsw(double x) {
  if (x <= 2.0) return 1.0;
  if (x >= 3.0) return 0.0;
  return (1.0 + cos(kPi * (x - 2.0))) / 2.0;
}
__device__ __forceinline__ double dFsw(double x) {
  if (x <= 2.0 || x >= 3.0) return 0.0;
  return (-kPi * sin(kPi * (x - 2.0))) / 2.0;
}
}  // namespace

template <bool VIR>
__global__ void __launch_bounds__(256) k_brenner_bonds(Atoms A, const Tables* __restrict__ T, Box bx,
                                                      const int2* __restrict__ bonds,
                                                      const int* __restrict__ nbonds, double* __restrict__ grad,
                                                      double* __restrict__ eatom, unsigned* err,
                                                      long long* __restrict__ vir) {
  const BrennerParams& P = T->bren;
  Virial<VIR> vw(vir, 1.0);
  const GradSink gsink{grad, A.gfix};
  int nbt = min(*nbonds, A.bond_cap);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nbt; t += gridDim.x * blockDim.x) {
    int2 bd = bonds[t];
    int i = bd.x, j = bd.y;
    if constexpr (VIR) vw.at(A.sys[i]);
    int ti = A.type[i], tj = A.type[j];
    D3 rv = depos(A.p, i, j, bx);
    double r = sqrt(norm2(rv));
    double f, df;
    cutoff(r, P.R[ti][tj][0], P.R[ti][tj][1], f, df);
    if (f == 0.0) continue;
    // VR / VA (Brenner.cxx:63-127)
    double S = P.S[ti][tj], De = P.De[ti][tj], be = P.beta[ti][tj], Re = P.Re[ti][tj];
    double aR = De / (S - 1.0), bR = sqrt(2.0 * S) * be;
    double aA = De * S / (S - 1.0), bA = sqrt(2.0 / S) * be;
    double vr = aR * exp(-bR * (r - Re));
    double va = aA * exp(-bA * (r - Re));
    double VR = f * vr, VA = f * va;
    double cr = vr * (-bR) * f + vr * df;  // VR seeded with 1
    double e = VR;
    D3 gi = d3(0, 0, 0), gj = d3(0, 0, 0);
    if (VA != 0.0) {
      double W = -VA;  // Baver seed
      // ---- forward
      double Dv[2], Bv[2], NHs[2], NCs[2], HN[2][3];
      for (int s = 0; s < 2; s++) {
        int a = s ? j : i, b = s ? i : j, ta = s ? tj : ti, tb = s ? ti : tj;
        D3 ra = s ? -rv : rv;
        double D = 1.0;
        int nb = A.bcnt[a];
        for (int q = 0; q < nb; q++) {
          BRef k = bref(A, a, q);
          if (k.k == b) continue;
          int tk = A.type[k.k];
          double al = P.alpha[ta][tb][tk];
          double E = (al == 0.0) ? 1.0 : exp(al * ((r - P.Re[ta][tb]) - (k.r - P.Re[ta][tk])));
          if (E == 0.0) continue;
          double G;
          if (ta == T_C) {
            double c = dot(ra, k.rv) / (r * k.r);
            if (!(c >= -1.0 - 1e-7 && c <= 1.0 + 1e-7)) raise(sys_err(err, A.sys[i]), ERR_COS_RANGE);
            if (c < -1.0) c = -1.0;
            if (c > 1.0) c = 1.0;
            G = P.ao * (1.0 + P.co * P.co / (P.do_ * P.do_) - P.co * P.co / (P.do_ * P.do_ + (1.0 + c) * (1.0 + c)));
          } else {
            G = P.G_HH;
          }
          D += G * k.f * E;
        }
        HN[s][0] = HN[s][1] = HN[s][2] = 0.0;
        NHs[s] = NCs[s] = 0.0;
        if (ta == T_C && tb <= T_H) {
          NHs[s] = nsum(A, a, b, T_H);
          NCs[s] = nsum(A, a, b, T_C);
          bicubic(T->BH[tb == T_C ? 0 : 1], NHs[s], NCs[s], HN[s]);
        }
        D += HN[s][0];
        Dv[s] = D;
        Bv[s] = pow(D, -P.delta[ta][ta]);
      }
      double FN[4] = {0, 0, 0, 0}, Ni = 0, Nj = 0, sums[2] = {0, 0}, Nconj = 1.0;
      if (ti == T_C && tj == T_C) {
        Ni = nsum(A, i, j, -1);
        Nj = nsum(A, j, i, -1);
        for (int s = 0; s < 2; s++) {
          int a = s ? j : i, b = s ? i : j;
          int nb = A.bcnt[a];
          for (int q = 0; q < nb; q++) {
            BRef k = bref(A, a, q);
            if (k.k == b || A.type[k.k] != T_C) continue;
            double x = nsum(A, k.k, a, -1);
            Nconj += k.f * Fsw(x);
          }
        }
        double fi = Ni, fj = Nj;
        double out[4];
        tricubic<double, 6>(T->BF, fi, fj, Nconj, 4, 5, out);
        for (int q = 0; q < 4; q++) FN[q] = out[q];
      }
      double bav = (Bv[0] + Bv[1]) / 2.0 + (FN[0] / 2.0);
      // ---- backward, seed W
      for (int s = 0; s < 2; s++) {
        int a = s ? j : i, b = s ? i : j, ta = s ? tj : ti, tb = s ? ti : tj;
        D3 ra = s ? -rv : rv;
        D3& ga = s ? gj : gi;
        D3& gb = s ? gi : gj;
        double dl = P.delta[ta][ta];
        double seedD = (-dl * pow(Dv[s], -dl - 1.0)) * (W / 2.0);
        int nb = A.bcnt[a];
        for (int q = 0; q < nb; q++) {
          BRef k = bref(A, a, q);
          if (k.k == b) continue;
          int tk = A.type[k.k];
          double al = P.alpha[ta][tb][tk];
          double E = (al == 0.0) ? 1.0 : exp(al * ((r - P.Re[ta][tb]) - (k.r - P.Re[ta][tk])));
          if (E == 0.0) continue;
          D3 gk = d3(0, 0, 0);
          double G;
          if (ta == T_C) {
            double raw = dot(ra, k.rv) / (r * k.r);
            double c = raw;
            if (c < -1.0) c = -1.0;
            if (c > 1.0) c = 1.0;
            G = P.ao * (1.0 + P.co * P.co / (P.do_ * P.do_) - P.co * P.co / (P.do_ * P.do_ + (1.0 + c) * (1.0 + c)));
            double dG = 2.0 * P.ao * (P.co * P.co) * (1.0 + c) /
                        ((P.do_ * P.do_ + (1.0 + c) * (1.0 + c)) * (P.do_ * P.do_ + (1.0 + c) * (1.0 + c)));
            double cc = dG * (k.f * E * seedD);
            double e2 = r * k.r;
            D3 da = k.rv / e2 - ra * (raw / (r * r));
            D3 db = ra / e2 - k.rv * (raw / (k.r * k.r));
            ga = ga + (da + db) * cc;
            gb = gb - da * cc;
            gk = gk - db * cc;
            vw.push(ra, da * cc);
            vw.push(k.rv, db * cc);
          } else {
            G = P.G_HH;
          }
          if (k.df != 0.0) {
            D3 v = k.rv * (k.df * (G * E * seedD) / k.r);
            ga = ga + v;
            gk = gk - v;
            vw.push(k.rv, v);
          }
          if (al != 0.0) {
            double dt = E * al;
            double V3 = G * k.f * seedD;
            D3 u = ra * (dt * V3 / r);
            ga = ga + u;
            gb = gb - u;
            D3 v = k.rv * (-dt * V3 / k.r);
            ga = ga + v;
            gk = gk - v;
            vw.push(ra, u);
            vw.push(k.rv, v);
          }
          atomic_add3(gsink, k.k, gk);
        }
        if (ta == T_C && tb <= T_H) {
          nsum_back(A, a, b, T_H, HN[s][1] * seedD, ga, grad, vw);
          nsum_back(A, a, b, T_C, HN[s][2] * seedD, ga, grad, vw);
        }
      }
      if (ti == T_C && tj == T_C) {
        double Wf = W / 2.0;
        nsum_back(A, i, j, -1, FN[1] * Wf, gi, grad, vw);
        nsum_back(A, j, i, -1, FN[2] * Wf, gj, grad, vw);
        double Vc = FN[3] * Wf;
        if (Vc != 0.0) {
          for (int s = 0; s < 2; s++) {
            int a = s ? j : i, b = s ? i : j;
            D3& ga = s ? gj : gi;
            int nb = A.bcnt[a];
            for (int q = 0; q < nb; q++) {
              BRef k = bref(A, a, q);
              if (k.k == b || A.type[k.k] != T_C) continue;
              double x = nsum(A, k.k, a, -1);
              if (k.df != 0.0) {
                D3 v = k.rv * (k.df * (Fsw(x) * Vc) / k.r);
                ga = ga + v;
                vw.push(k.rv, v);
                atomic_add3(gsink, k.k, -v);
              }
              double c2 = k.f * dFsw(x) * Vc;
              if (c2 != 0.0) {
                D3 gk = d3(0, 0, 0);
                nsum_back(A, k.k, a, -1, c2, gk, grad, vw);
                atomic_add3(gsink, k.k, gk);
              }
            }
          }
        }
      }
      // VA seeded with -b (Brenner.cxx:52-55)
      cr += va * (-bA) * f * (-bav) + va * (-bav) * df;
      e += -bav * VA;
    }
    D3 v = rv * (cr / r);
    gi = gi + v;
    gj = gj - v;
    vw.push(rv, v);
    atomic_add3(gsink, i, gi);
    atomic_add3(gsink, j, gj);
    energy_add(A, eatom, i, e);
  }
  vw.flush();
}

template __global__ void k_brenner_bonds<false>(Atoms, const Tables*, Box, const int2*, const int*, double*,
                                                double*, unsigned*, long long*);
template __global__ void k_brenner_bonds<true>(Atoms, const Tables*, Box, const int2*, const int*, double*,
                                               double*, unsigned*, long long*);

}  // namespace mdg
