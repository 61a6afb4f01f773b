// Kernel declarations and the device view of one system.
#pragma once
#include "mdg_device.hpp"

namespace mdg {

// One bond-list slot of atom a: the minimum-image vector x_a - x_k, r, the
// REBO cutoff f and df/dr, in one 32-byte (FP32) or 64-byte (FP64) record so
// that a bond is one contiguous load.
template <class F>
struct alignas(8 * sizeof(F)) BondRec {
  F x, y, z, r, f, df, pad0, pad1;
};

// Thermal bath of one system (SimLoop.hpp:118-189): friction
// -V M gamma (1 - sqrt(To/T)) on the atoms inside the geometry; lengths in cm.
enum : int { TB_NONE = 0, TB_UNIVERSE = 1, TB_BOX = 2, TB_SPHERE = 3 };
struct Bath {
  int geom;
  double To, gamma, zMin, dBoundary, zFree, cx, cy, cz, radius;
};

// Device view of the atoms and of the per-step derived data (SoA in HBM).
// A batch of independent systems is one array of atoms: system s owns atoms
// [sys_off[s], sys_off[s+1]) and sys[i] names atom i's system.
struct Atoms {
  int n;
  int nsys;
  int* sys;
  int* sys_off;
  double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;
  double4* p;                 // packed x, y, z and the type (as a double) for gathers; every position writer refreshes it
  double* grad;   // n*3 interleaved, dE/dx
  double *mass, *Z;
  int *type, *elem;
  unsigned char* fixed;
  int *pcx, *pcy, *pcz;       // PBC image counters (Atom::PBC_count)
  double *disp_s, *disp_l;    // n*3 displacement since the last rebuild
  int cap_s, cap_l;           // row strides of the short / long lists
  int *nls_cnt, *nls, *nll_cnt, *nll;
  int* nll_split;             // per long-list row: index of the first neighbour j > i
  // LJ pair rows: the j > i entries of each long-list row that lay within
  // the pair's LJ cutoff plus a margin when the row was built, in row order;
  // pl_cnt[i] = -1 when they did not fit cap_p (the LJ pass then sweeps the
  // long row).  Rebuilt with the long list and whenever the two largest
  // displacements since the last build (disp_p) add up to more than the margin.
  int cap_p;
  int *pl_cnt, *pl;
  double* disp_p;
  // Metal rows (embedded-atom members): the metal partners of each metal
  // atom within the members' largest range plus a skin when the row was
  // built, in long-row order; ml_cnt[i] = -1 when they did not fit
  // kMetalRow.  Rebuilt with the long list and when the two largest
  // displacements since the last build (disp_e) add up to more than the skin.
  int *ml_cnt, *ml;
  double* disp_e;
  int maxb;                   // bond-list capacity per atom
  int bond_cap;               // capacity of the (i<j) bond array
  int* bcnt;
  int* bj;
  BondRec<double>* b64;       // FP64 bond geometry (cm)
  BondRec<float>* b32;        // FP32 bond geometry (Angstrom), mixed precision
  double *Nt, *NH, *NC;       // REBO coordination sums (REBO.hpp:250-258)
  // Per atom a and partner type t (C, H): an upper bound of D_a^-1/2 over
  // every partner direction that is not a's bond (coordination pass; null
  // without AIREBO-LJ), the b* bound of the LJ pass.
  float* pmax;
  // thermal bath (null when no system of the context has one)
  const Bath* bath;           // per system
  const unsigned char* tb_apply;  // per atom: Atom::apply_ThermalBath
  double* tb_T;               // per system: bath temperature at the start of the step
  double* tb_dE;              // per atom: energy taken from the bath this step
  double* an_nt;              // n*3 acceleration without the bath (kept with or without a bath)
  int wg_map;                 // 1: the scattering force kernels give each XCD a contiguous eighth of the work
  // Deterministic accumulation (FP64 form; null in the mixed form): the
  // gradient and the atomically summed per-atom energies as fixed-point
  // pairs (hi, lo) -- gfix[6 a + 2 c], efix[2 (t n + a)] for the energy slot
  // eatom0 + t n -- joined into grad and eatom at the end of each force call.
  long long* gfix;
  long long* efix;
  double* eatom0;
  // Per atom: 1 when some long-range chain member lists a neighbour of it
  // (FProxy::hasNeighbors, FProxy.hpp:99-104, for those members), refreshed
  // at each long-list build; null without long-range members.
  unsigned char* nbl;
  // Mixed systems (atoms open on a periodic axis of their box, mdg_set_atom_pbc):
  // per atom the bits of those axes, null when no system is mixed; boxL = the
  // systems' boxes (3 per system).
  const unsigned char* open_axes;
  const double* boxL;
};

// The long-range chain members as FProxy::hasNeighbors sees them: member q
// lists atom i's partner j iff both elements are in handled[q] (bit per
// element index: H, C, Ar, Xe, Cu, Ag, Au) and d^2 < r2[q] = ((1 + 0.44)
// Rc_q)^2 (NeighbourList.cxx:83-102).
struct LongMembers {
  int m;
  unsigned handled[6];
  double r2[6];
};
__device__ __forceinline__ unsigned elem_bit(int e) {
  switch (e) {
    case 1: return 1u;
    case 12: return 2u;
    case 40: return 4u;
    case 131: return 8u;
    case 64: return 16u;
    case 108: return 32u;
    case 197: return 64u;
    default: return 0u;
  }
}
// flag: the long list's rebuild flag; sflags[2 s + 1]: system s rebuilt it.
__global__ void k_long_nbflag(const int* flag, const int* sflags, Atoms A, Box bx, LongMembers lm);

// Adds e (erg) to the per-atom energy slot[owner]: fixed point in the FP64
// form, an FP64 atomic otherwise.
__device__ __forceinline__ void energy_add(const Atoms& A, double* slot, int owner, double e) {
  if (e == 0.0) return;
  if (A.efix) fix_atomic(A.efix + 2 * ((slot - A.eatom0) + owner), e * kFixEnergy);
  else unsafeAtomicAdd(slot + owner, e);
}

struct CellGrid {
  double lo[3], w[3];
  int n[3];
  int periodic[3];
};

__global__ void k_cell_clear(const int* flag, const int* sflags, int list_id, const int* cell_off, int* cnt);
__global__ void k_cell_assign(const int* flag, const int* sflags, int list_id, int n, const double4* p,
                              const CellGrid* grids, const int* cell_off, const int* sys, int* cell_of,
                              int* cell_cnt);
__global__ void k_cell_scan(const int* flag, const int* sflags, int list_id, const int* cell_off, const int* sys_off,
                            int* cnt, int* start);
__global__ void k_cell_fill(const int* flag, const int* sflags, int list_id, int n, const int* sys,
                            const int* cell_of, const int* start, int* cursor, int* cell_atoms);
// APBC: the instantiation for the systems whose atoms carry their own PBC
// vectors (every pair of such a system is a candidate, tested with depos of
// the true coordinates xs/ys/zs); the other one skips those systems.
template <int CAP, bool SORT, bool APBC>
__global__ void k_nl_build(const int* flag, const int* sflags, int list_id, int n, const double4* p,
                           const unsigned char* handled, Box bx, const CellGrid* grids,
                           const int* cell_off, double range2, const int* cell_of, const int* start,
                           const int* cell_atoms, int* cnt, int* list, double* disp, unsigned* err, int* split,
                           const double* xs, const double* ys, const double* zs);
// Per-system reductions are tiled: tile t = rb[t] = {system, first atom},
// at most kRedTile atoms, one 256-thread block each, writing 8 doubles to
// part[8 t ..]; a 64-thread block per system then merges the tiles
// rb_off[s] .. rb_off[s+1] in order.
constexpr int kRedTile = 512;
// flags[0..2]: some system rebuilds its short list / long list / LJ pair
// rows; sflags[2*s+{0,1}] and pflags[s] per system.
__global__ void k_sys_disp_part(const int2* rb, const int* sys_off, const double* ds, const double* dl,
                                const double* dp, const double* de, double* part);
// flags[3] / eflags[s]: metal rows to rebuild (with the long list, or past lim_e)
__global__ void k_sys_disp_merge(const int* rb_off, const double* part, double lim_s, double lim_l, double lim_p,
                                 int have_p, double lim_e, int have_e, int* sflags, int* pflags, int* eflags,
                                 int* flags, long long* rebuilds);
__global__ void k_clear_flags(int nsys, int* flags, int* sflags, int* pflags, int* eflags);
// Metal rows of the systems flagged in eflags (any flagged: *flag), partners
// of type T_OTHER within rcand; 16 lanes per atom, blocks of 256.
constexpr int kMetalRow = 128;
__global__ void k_metal_rows(Atoms A, Box bx, const int* flag, const int* eflags, double rcand);
__global__ void k_bond_enum(Atoms A, int2* out, int* count, int cap, unsigned* err, long long* ntrip);
// Energies go to eatom (per-atom, per chain member), owned by the first atom
// of the bond, pair or quadruplet.
template <class R>
__global__ void k_coord(Atoms A, const TablesT<R>* T, Box bx, unsigned* err);
// TH = true (FP64 only): also accumulate dE/dtheta per system into dth
// (kNumTheta doubles per system).  OCC: waves per SIMD the bond-order
// kernels' register allocation is held to (1 or 2; chosen by the autotuner).
// REBO bonds with the shells of each wavefront's tile of 64 bonds staged in
// LDS (FP32 and FP64 forms, no parameter gradients or virial).
template <class R, int OCC>
__global__ void k_rebo_shell(Atoms A, const TablesT<R>* T, const Tables* T64, Box bx, const int2* bonds,
                             const int* nbonds, double* grad, double* eatom, unsigned* err);
template <class R, bool TH, int OCC, bool VIR>
__global__ void k_rebo_bonds(Atoms A, const TablesT<R>* T, const Tables* T64, Box bx, const int2* bonds,
                             const int* nbonds, double* grad, double* eatom, unsigned* err, double* dth,
                             long long* vir);
// The b* work queue is kQueues sub-queues of qsub entries each; sub-queue q
// holds entries [q * qsub, (q + 1) * qsub) and its count sits at
// qcount[q * kQStride], one count per 128-byte line, so the LJ kernel's
// appends spread over many addresses instead of serialising on one.
constexpr int kQueues = 256;
constexpr int kQStride = 32;
// Event tallies (LJ pairs, torsion quads, pair rows built) are spread over
// kTallySlots slots, one per 128-byte line, picked by workgroup: tally kind k
// of slot q is tally[q * kTallyStride + k]; the host adds the slots up.
constexpr int kTallySlots = 256;
constexpr int kTallyStride = 16;

// Adds a per-lane tally into this workgroup's slot of a spread tally (kind
// pointer = tally + k) with one atomic per wave (all 64 lanes must be
// active): one address bumped by every wave of a grid would serialise it.
__device__ __forceinline__ void count_add(long long* kind, long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v)
    atomicAdd((unsigned long long*)(kind + (blockIdx.x & (kTallySlots - 1)) * kTallyStride), (unsigned long long)v);
}

template <class R>
__global__ void k_pair_rows(Atoms A, const TablesT<R>* T, Box bx, const int* pflag, const int* pflags,
                            double margin, long long* rebuilt);
// VIR (template flag of every force kernel): also add the per-system virial
// of the kernel's terms into vir[6 s + c] (xx, yy, zz, xy, xz, yz; erg).
template <class R, bool TH, int OCC, int LW, bool VIR>
__global__ void k_lj_wave(Atoms A, const TablesT<R>* T, Box bx, double* grad, double* eatom, int4* queue,
                          int* qcount, int qsub, long long* npairs, unsigned* err, double* dth, long long* vir);
template <class R, bool TH, int OCC, int PASS, bool VIR>
__global__ void k_bstar(Atoms A, const TablesT<R>* T, Box bx, int4* queue, const int* qcount, int qsub,
                        double* grad, double* eatom, unsigned* err, double* dth, long long* vir,
                        long long* work);
template <class R, bool VIR>
__global__ void k_etors(Atoms A, const TablesT<R>* T, Box bx, const int2* bonds, const int* nbonds, int compat,
                        double* grad, double* eatom, long long* nquads, long long* vir);
template <bool VIR>
__global__ void k_fbzl(Atoms A, Box bx, const int* ions, int nions, double R0, double R1, double* grad, double* eatom,
                       long long* vir);
template <bool VIR>
__global__ void k_flj(Atoms A, Box bx, const FLJTable* T, double* grad, double* eatom, long long* vir);
// Ackland EAM: the density pass (F'(rho) into fe, F into eatom), then the
// force pass (gradient and pair energies); 16 lanes per atom, blocks of 256.
// ek: per atom the EAM element index (Cu 0, Ag 1, Au 2) or -1; fe: per atom
// (F'(rho), element index) written by the density pass.
// row/rcnt: per atom its partners inside the EAM range this step (kEamRow
// slots each), written by the density pass and read by the force pass.
constexpr int kEamRow = 128;
// The two embedded-atom forms, as policies of the two passes: the element
// pair's density function g and slope, its pair term V and slope, and the
// embedding F(rho) with F'(rho) (0 where rho is 0: F() applies no gradient
// there).  Members are defined with the kernels.
struct AcklandModel {  // Ackland.cxx:59-221
  using Table = EAMTable;
  static __device__ double rmax(const Table* T);
  static __device__ void g(const Table* T, int a, int b, double r, double& g, double& dg);
  static __device__ void pair(const Table* T, int a, int b, double r, double& v, double& dv);
  static __device__ void embed(const Table* T, double rho, double& F, double& Fp);
};
struct TBModel {  // TightBinding.cxx:59-159
  using Table = TBTable;
  static __device__ double rmax(const Table* T);
  static __device__ void g(const Table* T, int a, int b, double r, double& g, double& dg);
  static __device__ void pair(const Table* T, int a, int b, double r, double& v, double& dv);
  static __device__ void embed(const Table* T, double rho, double& F, double& Fp);
};
// The density pass sweeps the metal rows (the long row where ml_cnt < 0).
template <class M>
__global__ void k_eam_density(Atoms A, Box bx, const typename M::Table* T, const signed char* ek, double2* fe,
                              double* eatom, int* row, int* rcnt, unsigned* err);
template <class M, bool VIR>
__global__ void k_eam_force(Atoms A, Box bx, const typename M::Table* T, const signed char* ek, const double2* fe,
                            const int* row, const int* rcnt, double* grad, double* eatom, long long* vir);
// FBM pairs (Cu-Cu, Cu-Ar), 16 lanes per Cu or Ar atom, blocks of 256.
template <bool VIR>
__global__ void k_fbm(Atoms A, Box bx, const FBMTable* T, double* grad, double* eatom, long long* vir);
template <bool VIR>
__global__ void k_brenner_bonds(Atoms A, const Tables* T, Box bx, const int2* bonds, const int* nbonds,
                                double* grad, double* eatom, unsigned* err, long long* vir);
// Verlet halves over the reduction tiles: the drift writes the skin test's
// tile part (as k_sys_disp_part), the kick the step outputs' tile part
// (merged by k_sys_step_merge).
// simt/tfin: per system the simulated time and the final time; a system
// with simt >= tfin stands still (no drift, no kick, dt and time kept).
__global__ void k_drift_tile(Atoms A, const int2* rb, const double* dtp, const double* simt, const double* tfin,
                             double* part);
__global__ void k_kick_tile(Atoms A, Box bx, const int2* rb, const double* dtp, const double* simt,
                            const double* tfin, int nterms, const double* eatom, double* part);
__global__ void k_init_acc(Atoms A, const unsigned char* keep);
__global__ void k_pack_pos(Atoms A);
__global__ void k_sys_energy(int nterms, const int* sys_off, int n, const double* eatom, double* eacc);
// serr[s]: system s's device error word; a system with an error stops (its
// final time is set to its current time).
__global__ void k_sys_step_merge(const int* rb_off, const double* part, double* dtp, int fixed_dt, double* out,
                                 const int* stepidx, int nsys, double* ebath, double* simt, double* tfin,
                                 double* dtprev, long long* iter, const unsigned* serr);
__global__ void k_bath_temp_part(Atoms A, Box bx, const int2* rb, double* part);
__global__ void k_bath_temp_merge(Atoms A, const int* rb_off, const double* part);
__global__ void k_fix_join(Atoms A, int nterms, unsigned fixmask, double* eatom);
__global__ void k_step_virial(int nsys, const double* vir, double* stepvir, const int* stepidx);
__global__ void k_vir_join(int nsys, const long long* fix, double* vir);
__global__ void k_step_advance(int* stepidx);

// Trajectory output captured inside the step (the layouts of mdg_frame and
// mdg_atom_record in mdg.h): at the start of an iteration a system due for a
// SnapshotList record (mdtrajsim's CustomSimLoop cadence) or an XVA frame
// (SimLoop::writetrajXVA cadence) reserves a frame and its atom records.
enum : int { TRAJ_SNAPSHOT = 1, TRAJ_XVA = 2 };
struct TrajRec {
  double x[3], v[3];
  int pc[3];
  int atom;
};
struct TrajFrame {
  int kind, system;
  long long iteration;
  double sim_time, dt;
  long long first, count;
};
struct TrajBuf {
  TrajRec* rec;
  TrajFrame* frm;
  unsigned long long* cnt;  // [0] frames, [1] records, [2] overflow
  long long cap_rec, cap_frm;
  const int* sel;      // selected atoms (SnapshotList), per system sel[sel_off[s] ..]
  const int* sel_off;  // nsys + 1, or null: no SnapshotList
  int xva;             // XVA frames wanted
};
__global__ void k_traj_capture(Atoms A, TrajBuf tb, const double* simt, const double* tfin, const double* dtprev,
                               const long long* iter, const double* dtp);
__global__ void k_zero(double* p, size_t n);
__global__ void k_zero_i(int* p, int n);

void build_tables(Tables& T, int brenner_set);
void convert_tables(const Tables& d, TablesT<float>& f);

}  // namespace mdg
