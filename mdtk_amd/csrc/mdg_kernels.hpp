// Kernel declarations and the device view of one system.
#pragma once
#include "mdg_device.hpp"

namespace mdg {

// Device view of the atoms and of the per-step derived data (SoA in HBM).
struct Atoms {
  int n;
  double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;
  double* grad;   // n*3 interleaved, dE/dx
  double *mass, *Z;
  int *type, *elem;
  unsigned char* fixed;
  int *pcx, *pcy, *pcz;       // PBC image counters (Atom::PBC_count)
  double *disp_s, *disp_l;    // n*3 displacement since the last rebuild
  int cap_s, cap_l;           // row strides of the short / long lists
  int *nls_cnt, *nls, *nll_cnt, *nll;
  int maxb;                   // bond-list capacity per atom
  int bond_cap;               // capacity of the (i<j) bond array
  int* bcnt;
  int* bj;
  double *br, *bf, *bdf, *bx, *by, *bz;       // FP64 geometry (cm)
  float *brf, *bff, *bdff, *bxf, *byf, *bzf;  // FP32 geometry (Angstrom), mixed precision
  double *Nt, *NH, *NC;       // REBO coordination sums (REBO.hpp:250-258)
};

struct CellGrid {
  double lo[3], w[3];
  int n[3];
  int periodic[3];
};

__global__ void k_cell_clear(const int* flag, int ncell, int* cnt);
__global__ void k_cell_assign(const int* flag, int n, const double* x, const double* y, const double* z, CellGrid g,
                              int* cell_of, int* cell_cnt);
__global__ void k_cell_scan(const int* flag, int ncell, int* cnt, int* start);
__global__ void k_cell_fill(const int* flag, int n, const int* cell_of, const int* start, int* cursor,
                            int* cell_atoms);
template <int CAP>
__global__ void k_nl_build(const int* flag, int n, const double* x, const double* y, const double* z,
                           const unsigned char* handled, Box bx, CellGrid g, double range2, const int* cell_of,
                           const int* start, const int* cell_atoms, int* cnt, int* list, double* disp,
                           unsigned* err);
__global__ void k_disp_top2(int n, const double* ds, const double* dl, double* part);
__global__ void k_disp_decide(int nblk, const double* part, double lim_s, double lim_l, int* flags,
                              long long* rebuilds);
__global__ void k_bond_enum(Atoms A, int2* out, int* count, int cap, unsigned* err);
template <class R>
__global__ void k_coord(Atoms A, const TablesT<R>* T, Box bx, unsigned* err);
template <class R>
__global__ void k_rebo_bonds(Atoms A, const TablesT<R>* T, Box bx, const int2* bonds, const int* nbonds,
                             double* grad, double* eacc, unsigned* err);
template <class R>
__global__ void k_lj_wave(Atoms A, const TablesT<R>* T, Box bx, double* grad, double* eacc, int4* queue,
                          int* qcount, int qcap, long long* npairs, unsigned* err);
template <class R>
__global__ void k_bstar(Atoms A, const TablesT<R>* T, Box bx, const int4* queue, const int* qcount, int qcap,
                        double* grad, double* eacc, unsigned* err);
template <class R>
__global__ void k_etors(Atoms A, const TablesT<R>* T, Box bx, const int2* bonds, const int* nbonds, int compat,
                        double* grad, double* eacc, long long* nquads);
__global__ void k_fbzl(Atoms A, Box bx, double R0, double R1, double* grad, double* eacc);
__global__ void k_brenner_bonds(Atoms A, const Tables* T, Box bx, const int2* bonds, const int* nbonds,
                                double* grad, double* eacc, unsigned* err);
__global__ void k_drift(Atoms A, const double* dtp);
__global__ void k_kick(Atoms A, Box bx, const double* dtp, double* red);
__global__ void k_init_acc(Atoms A);
__global__ void k_step_end(int nterms, double* eacc, double* red, double* dtp, int fixed_dt, double* out, int* stepidx);
__global__ void k_zero(double* p, size_t n);
__global__ void k_zero_i(int* p, int n);

void build_tables(Tables& T, int brenner_set);
void convert_tables(const Tables& d, TablesT<float>& f);

}  // namespace mdg
