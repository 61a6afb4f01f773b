// Host side of the C ABI declared in include/mdg.h: one context owns the
// device buffers of one system or of a batch of independent systems, the
// parameter tables and a HIP stream; every call is stream-ordered and
// synchronous unless it says otherwise.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mdg.h"
#include "mdg_kernels.hpp"

using namespace mdg;

namespace {
thread_local std::string g_create_err;

struct KStat {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0;
  long long launches = 0;
};

// Launch shapes of the force kernels (0 = the built-in heuristic); set by
// mdg_autotune from timings on the context's own system.
struct Tune {
  // waves/SIMD variant of the bond-order kernels; occ_bonds 11 and 12: the
  // REBO bonds with LDS-staged shells (k_rebo_shell) at 1 or 2 waves/SIMD
  int occ_bonds = 1, occ_bstar = 1;
  // LJ kernel shape: 0 = 64 lanes per atom, 1 = the same held to 3 waves/SIMD,
  // 2 = 32 lanes per atom, 3 = 16 lanes per atom (more atoms in flight)
  int lj_shape = 3;
  int grid_bonds = 0, grid_bstar = 0, grid_lj = 0, grid_tors = 0;
};
}  // namespace

constexpr int kGraphBlock = 32;  // steps per launch of the multi-step graph

struct mdg_ctx {
  mdg_config cfg{};
  std::string err;
  std::vector<int> kinds;
  int slot_fbzl = -1, slot_flj = -1, slot_rebo = -1, slot_bren = -1, slot_lj = -1, slot_tors = -1, slot_eam = -1,
      slot_fbm = -1, slot_tb = -1;
  FLJTable* dFLJ = nullptr;  // FLJ parameters (device)
  EAMTable* dEAM = nullptr;  // Ackland parameters (device)
  int2* rb = nullptr;          // reduction tiles {system, first atom}
  int* rb_off = nullptr;       // per system: its first tile (nsys + 1)
  double* red_part = nullptr;  // 8 partial results per tile
  double* simt = nullptr;      // per system: simulated time (s)
  double* tfin = nullptr;      // per system: final time (s); a system past it stands still
  double* dtprev = nullptr;    // per system: dt of the previous iteration (SimLoop::dt_prev)
  long long* iter = nullptr;   // per system: iterations done (SimLoop::iteration)
  // trajectory output (mdg_set_trajectory_output)
  bool traj_on = false;
  TrajBuf tb{};
  std::vector<mdg_frame> frames_h;
  std::vector<mdg_atom_record> recs_h;
  int nred = 0;
  // per embedded-atom member (0 Ackland, 1 TightBinding): per atom its element
  // index or -1, (F'(rho), element) of the density pass, and its partners
  // inside the range this step (kEamRow slots + count)
  struct EmbedSet {
    signed char* e = nullptr;
    double2* fe = nullptr;
    int* row = nullptr;
    int* cnt = nullptr;
  } emb[2];
  TBTable* dTB = nullptr;    // TightBinding parameters (device)
  FBMTable* dFBM = nullptr;  // FBM parameters (device)
  bool have_short = false, have_long = false;
  LongMembers lmem{};  // the long-range members for the dt rule's hasNeighbors
  double rc_short = 0, rc_long = 0;

  int n = 0;
  int nsys = 0;
  std::vector<double> Lh;      // 3*nsys box lengths (host copy)
  std::vector<int> off_h;      // nsys+1 atom offsets (host copy)
  std::vector<double> dt_host;  // per-system dt carried between calls
  Box box{};                   // device: L[3*nsys], sys[n]; per-atom PBC of mixed systems
  // Systems whose atoms carry PBC vectors other than their box (mdg_set_atom_pbc)
  // Systems whose atoms carry their own PBC vectors (mdg_set_atom_pbc): per
  // system a flag, per atom the periodic axes of the box on which the atom
  // is open (host copies; empty: no such system)
  std::vector<unsigned char> apbc_h, open_h;
  std::vector<unsigned char> tb_apply_h;  // per atom: Atom::apply_ThermalBath (host copy; empty: all)
  unsigned char *apbc_d = nullptr, *open_d = nullptr;
  Atoms A{};
  Tables* dT = nullptr;
  TablesT<float>* dTf = nullptr;  // mixed precision: Angstrom/eV FP32 table
  bool mixed = false;
  hipStream_t own = nullptr, stream = nullptr;
  // second stream (default priority) for the bond-centred force kernels, which
  // run beside the LJ chain; joined back into `stream` by events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<void*> allocs;
  // With arena set, dalloc carves device buffers from [arena, arena +
  // arena_cap) instead of calling hipMalloc (mdg_energy_grad_noalloc).
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;

  // one cell grid per system and list; cell ids offset per system
  CellGrid *gs_d = nullptr, *gl_d = nullptr;
  int *coff_s = nullptr, *coff_l = nullptr;
  int ncs = 0, ncl = 0;
  int *cell_of_s = nullptr, *cnt_s = nullptr, *start_s = nullptr, *atoms_s = nullptr;
  int *cell_of_l = nullptr, *cnt_l = nullptr, *start_l = nullptr, *atoms_l = nullptr;
  unsigned char *handled_s = nullptr, *handled_l = nullptr;
  int* flags = nullptr;   // [0] short, [1] long list, [2] LJ pair rows, [3] metal rows: rebuild requested by some system
  int* sflags = nullptr;  // [2*s + list] rebuild requested by system s
  int* pflags = nullptr;  // [s] LJ pair rows of system s to be rebuilt
  double pair_margin = 1.0e-8;  // LJ pair-row margin (cm)
  int* eflags = nullptr;  // [s] metal rows of system s to be rebuilt
  double metal_skin = 1.0e-8;   // metal-row skin (cm)
  double metal_rcand = 0.0;     // metal-row radius: the embedded-atom members' largest range + skin
  int2* bonds = nullptr;
  int* nbonds = nullptr;
  int bond_cap = 0;
  int4* queue = nullptr;
  int* qcount = nullptr;
  int qsub = 0;  // entries per b* sub-queue (kQueues of them)
  double* eatom = nullptr;  // nterms*n per-atom energies of the last force call
  bool virial = false;      // force calls reduce the virial (mdg_set_virial)
  double* vir = nullptr;    // 6*nsys virial of the last force call (erg)
  long long* virfix = nullptr;  // 12*nsys: its six components as fixed-point (hi, lo) sums
  double* stepvir = nullptr;  // chunk*nsys*6 per-step virials of mdg_step_nve
  int stepvir_cap = 0;
  std::vector<double> vir_steps;  // per-step virials of the last mdg_step_nve
  double* dtheta = nullptr;  // kNumTheta*nsys parameter gradients (allocated on first use)
  bool theta_mode = false;
  Tables host_T;             // parameter tables as uploaded (for parameter overrides)
  double* eacc = nullptr;   // 8*nsys per-system energies of each chain member
  double* dtp = nullptr;    // nsys
  double* stepout = nullptr;
  int step_cap = 0;         // steps per output chunk
  long long* counters = nullptr;  // [0] short rebuilds [1] long rebuilds (per system and rebuild)
  // kTallySlots spread slots: [0] LJ pairs [1] quads [2] pair rows built [3] REBO sigma-pi triplets
  // [4] b* bound-pass sigma-pi triplets [5] b* pairs given the full bond order
  long long* tally = nullptr;
  unsigned* derr = nullptr;  // device error words: [0] global resources, [1 + s] system s
  int* ions = nullptr;       // the atoms the ZBL term centres on (Ar, Xe), ascending
  int nions = 0;
  std::vector<unsigned> failed;  // per system: error bits seen so far (mdg_system_errors)
  long iteration = 0;

  Tune tune;
  bool profiling = false;
  unsigned char* acc_keep = nullptr;  // per system: 1 = keep the accelerations set (a resumed state)
  std::map<std::string, KStat> kst;
  hipEvent_t timer_ev[8] = {};
  std::vector<hipEvent_t> event_pool;

  int* stepidx = nullptr;
  // thermal bath (allocated on first use)
  std::vector<Bath> bath_h;
  Bath* bath_d = nullptr;
  unsigned char* tb_apply_d = nullptr;
  double* ebath = nullptr;  // per system: energy taken from the bath so far (erg)
  bool bath_on = false;     // some system has a bath geometry
  bool use_graphs = true;
  int graph_block = kGraphBlock;  // steps per captured multi-step graph (1: one step per launch)
  // [fixed + 2 * block]: one step, or kGraphBlock steps in one graph, with
  // adaptive / fixed dt
  hipGraphExec_t graph_exec[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<hipGraph_t> graphs;

  ~mdg_ctx() {
    if (stream) hipStreamSynchronize(stream);
    for (auto& e : graph_exec)
      if (e) hipGraphExecDestroy(e);
    for (auto& g : graphs) hipGraphDestroy(g);
    for (auto& kv : kst)
      for (auto& pr : kv.second.pending) {
        hipEventDestroy(pr.first);
        hipEventDestroy(pr.second);
      }
    for (auto& e : timer_ev)
      if (e) hipEventDestroy(e);
    for (void* p : allocs) hipFree(p);  // the tables too; nothing when the buffers came from an arena
    if (own) hipStreamDestroy(own);
    if (side) hipStreamDestroy(side);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_join) hipEventDestroy(ev_join);
  }
};

namespace {
#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

template <class T>
int dalloc(mdg_ctx* c, T*& p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (c->arena) {  // carve from the caller's memory, 256-byte aligned
    size_t at = (c->arena_used + 255) & ~size_t(255);
    if (at + bytes > c->arena_cap) {
      c->err = "static scratch exhausted: this system needs more than its " + std::to_string(c->arena_cap >> 20) +
               " MiB";
      return 1;
    }
    p = reinterpret_cast<T*>(c->arena + at);
    c->arena_used = at + bytes;
    return 0;
  }
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    c->err = "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e);
    return 1;
  }
  c->allocs.push_back(q);
  p = static_cast<T*>(q);
  return 0;
}

// Frees a buffer that dalloc made (and forgets it); null is a no-op.
template <class T>
void dfree(mdg_ctx* c, T*& p) {
  if (!p) return;
  void* q = const_cast<void*>(static_cast<const void*>(p));
  auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
  if (it != c->allocs.end()) {
    hipFree(q);
    c->allocs.erase(it);
  }
  p = nullptr;
}

int type_of(int elem) {
  switch (elem) {
    case 12: return T_C;
    case 1: return T_H;
    case 40:
    case 131: return T_ION;
    default: return T_OTHER;
  }
}

double mass_of(int elem) {  // mdtk/Atom.hpp:98-113
  switch (elem) {
    case 1: return 1.0 * kAmu;
    case 12: return 12.0 * kAmu;
    case 64: return 63.546 * kAmu;
    case 108: return 107.868 * kAmu;
    case 197: return 196.967 * kAmu;
    case 40: return 39.948 * kAmu;
    case 131: return 131.293 * kAmu;
    default: return -1.0;
  }
}

double charge_of(int elem) {
  switch (elem) {
    case 1: return 1.0 * kECharge;
    case 12: return 6.0 * kECharge;
    case 64: return 29.0 * kECharge;
    case 108: return 47.0 * kECharge;
    case 197: return 79.0 * kECharge;
    case 40: return 18.0 * kECharge;
    case 131: return 54.0 * kECharge;
    default: return 0.0;
  }
}

// Launch wrapper: records HIP events around the launch when profiling.
template <class F>
void launch(mdg_ctx* c, const char* name, F&& f) {
  if (!c->profiling) {
    f();
    return;
  }
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, c->stream);
  f();
  hipEventRecord(b, c->stream);
  c->kst[name].pending.push_back({a, b});
}

void collect_profile(mdg_ctx* c) {
  for (auto& kv : c->kst) {
    for (auto& pr : kv.second.pending) {
      hipEventSynchronize(pr.second);
      float ms = 0;
      hipEventElapsedTime(&ms, pr.first, pr.second);
      kv.second.ms += ms;
      kv.second.launches += 1;
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
    kv.second.pending.clear();
  }
}

CellGrid make_grid(const double L[3], const std::vector<double>& xyz, int n, double range) {
  CellGrid g{};
  for (int d = 0; d < 3; d++) {
    if (L[d] != kNoPbc) {
      g.periodic[d] = 1;
      g.lo[d] = 0.0;
      int m = (int)std::floor(L[d] / range);
      if (m < 1) m = 1;
      g.n[d] = m;
      g.w[d] = L[d] / m;
    } else {
      g.periodic[d] = 0;
      double lo = 1e300, hi = -1e300;
      for (int i = 0; i < n; i++) {
        lo = std::min(lo, xyz[3 * i + d]);
        hi = std::max(hi, xyz[3 * i + d]);
      }
      double margin = 4.0 * range;
      lo -= margin;
      hi += margin;
      double ext = hi - lo;
      int m = (int)std::floor(ext / range);
      if (m < 1) m = 1;
      if (m > 512) m = 512;
      g.n[d] = m;
      g.lo[d] = lo;
      g.w[d] = ext / m;
    }
  }
  return g;
}

// The cell grids of every system for both lists, from host positions p
// (3n): a system whose atoms carry their own PBC vectors gets one cell, so
// its list build tests every pair of the system with depos, as the
// reference's all-pairs build does (NeighbourList.cxx:31-106).
int setup_grids(mdg_ctx* c, const std::vector<double>& p) {
  const int nsys = c->nsys;
  const std::vector<int>& off = c->off_h;
  auto one = [&](double rc, CellGrid*& gd, int*& coff, int& ncell, int*& cnt, int*& start) -> int {
    std::vector<CellGrid> g(nsys);
    std::vector<int> co(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
      if (!c->apbc_h.empty() && c->apbc_h[s]) {
        g[s] = CellGrid{};
        for (int d = 0; d < 3; d++) {
          g[s].n[d] = 1;
          g[s].w[d] = 1e300;
        }
      } else {
        std::vector<double> ps(p.begin() + 3 * (size_t)off[s], p.begin() + 3 * (size_t)off[s + 1]);
        g[s] = make_grid(&c->Lh[3 * s], ps, off[s + 1] - off[s], (1.0 + 0.44) * rc);
      }
      long long cells = (long long)g[s].n[0] * g[s].n[1] * g[s].n[2];
      if (co[s] + cells > (1LL << 28)) {
        c->err = "cell grid too large";
        return 1;
      }
      co[s + 1] = co[s] + (int)cells;
    }
    ncell = co[nsys];
    dfree(c, gd);
    dfree(c, coff);
    dfree(c, cnt);
    dfree(c, start);
    if (dalloc(c, gd, nsys) || dalloc(c, coff, nsys + 1) || dalloc(c, cnt, ncell) || dalloc(c, start, ncell + 1))
      return 1;
    HIPCHK(c, hipMemcpy(gd, g.data(), nsys * sizeof(CellGrid), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(coff, co.data(), (nsys + 1) * sizeof(int), hipMemcpyHostToDevice));
    return 0;
  };
  if (c->have_short && one(c->rc_short, c->gs_d, c->coff_s, c->ncs, c->cnt_s, c->start_s)) return 1;
  if (c->have_long && one(c->rc_long, c->gl_d, c->coff_l, c->ncl, c->cnt_l, c->start_l)) return 1;
  return 0;
}

std::string err_text(unsigned h) {
  std::string m;
  if (h & ERR_SHORT_NL_OVERFLOW) m += " short-list overflow (raise max_short_nb);";
  if (h & ERR_LONG_NL_OVERFLOW) m += " long-list overflow (raise max_long_nb);";
  if (h & ERR_BOND_OVERFLOW) m += " more than 16 bonds on one atom;";
  if (h & ERR_COS_RANGE) m += " cos(theta) outside [-1-1e-7, 1+1e-7] (FGeneral.hpp:147);";
  if (h & ERR_NEG_COORD) m += " coordination below -1e-15 (REBO.cxx:434);";
  if (h & ERR_QUEUE_OVERFLOW) m += " work-queue overflow;";
  if (h & ERR_SPLINE_DOMAIN) m += " spline argument outside its domain;";
  if (h & ERR_EAM_ROW_OVERFLOW) m += " more than 128 metal atoms inside the Ackland range of one atom;";
  if (h & ERR_PATH_OVERFLOW) m += " an atom's LJ pairs or C_ij paths beyond their capacity (128 pairs, 32 two-bond "
                                  "prefixes, 128 path records, 64 targets);";
  if (h & ERR_TYPE_RANGE) m += " a packed position carries a type outside the tables;";
  return m;
}

// Reads the device error words: w[0] (global resources) and w[1 + s] (system
// s).  Errors of global resources always fail the call.  Errors of single
// systems fail it too, unless keep_going: then the systems are recorded in
// c->failed (their words stay set, so the step loop keeps them still) and the
// call carries on.  Returns 1 on failure (message in c->err).
int check_err(mdg_ctx* c, bool keep_going = false) {
  const int ns = std::max(c->nsys, 0);
  std::vector<unsigned> h(1 + (size_t)ns, 0u);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->derr, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  unsigned any = h[0];
  for (int s = 0; s < ns; s++) any |= h[1 + s];
  if (!any) return 0;
  std::string m = "device error:";
  if (h[0]) m += err_text(h[0]);
  for (int s = 0; s < ns; s++)
    if (h[1 + s]) {
      if (ns > 1) m += " [system " + std::to_string(s) + "]";
      m += err_text(h[1 + s]);
      c->failed[s] |= h[1 + s];
    }
  if (!h[0] && keep_going) {
    c->err = m;
    return 0;
  }
  std::vector<unsigned> z(h.size(), 0u);  // a failed call leaves the words clear for the next
  HIPCHK(c, hipMemcpy(c->derr, z.data(), z.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  c->err = m;
  return 1;
}

inline int blocks(long long work, int bs = 256, int cap = 1 << 20) {
  long long b = (work + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

// Entries in the b* sub-queues after the last force call.
hipError_t queued(mdg_ctx* c, int& nq) {
  std::vector<int> h((size_t)kQueues * kQStride);
  hipError_t e = hipMemcpy(h.data(), c->qcount, h.size() * sizeof(int), hipMemcpyDeviceToHost);
  nq = 0;
  for (int q = 0; q < kQueues; q++) nq += std::min(h[(size_t)q * kQStride], c->qsub);
  return e;
}

// Requests a rebuild of every system's LJ pair rows and metal rows at the
// next force call (the rows are filtered from a long list just rebuilt).
int mark_pair_rows(mdg_ctx* c) {
  if (c->nsys == 0 || (!c->A.pl && !c->A.ml)) return 0;
  std::vector<int> one((size_t)c->nsys, 1);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->A.pl) {
    HIPCHK(c, hipMemcpy(c->pflags, one.data(), one.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->flags + 2, one.data(), 4, hipMemcpyHostToDevice));
  }
  if (c->A.ml) {
    HIPCHK(c, hipMemcpy(c->eflags, one.data(), one.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->flags + 3, one.data(), 4, hipMemcpyHostToDevice));
  }
  return 0;
}

// Allocates the bath's device state of the current systems (no geometry,
// every atom opted in, nothing transferred yet).
int bath_alloc(mdg_ctx* c) {
  if (c->bath_d) return 0;
  Atoms& A = c->A;
  int n = c->n, nsys = c->nsys;
  double* tbT = nullptr;
  double* tbdE = nullptr;
  if (dalloc(c, c->bath_d, nsys) || dalloc(c, c->tb_apply_d, n) || dalloc(c, tbT, nsys) || dalloc(c, tbdE, n) ||
      dalloc(c, c->ebath, nsys))
    return 1;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemset(c->tb_apply_d, 1, n));
  HIPCHK(c, hipMemset(tbT, 0, nsys * sizeof(double)));
  HIPCHK(c, hipMemset(tbdE, 0, n * sizeof(double)));
  HIPCHK(c, hipMemset(c->ebath, 0, nsys * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->bath_d, c->bath_h.data(), nsys * sizeof(Bath), hipMemcpyHostToDevice));
  A.bath = c->bath_d;
  A.tb_apply = c->tb_apply_d;
  A.tb_T = tbT;
  A.tb_dE = tbdE;
  return 0;
}

void drop_graphs(mdg_ctx* c) {  // the step's launches changed
  for (auto& e : c->graph_exec)
    if (e) {
      hipGraphExecDestroy(e);
      e = nullptr;
    }
}

// Neighbour-list rebuild for one list, gated on the device flag.
void nl_build(mdg_ctx* c, bool shortlist) {
  const int* flag = c->flags + (shortlist ? 0 : 1);
  const CellGrid* g = shortlist ? c->gs_d : c->gl_d;
  const int* coff = shortlist ? c->coff_s : c->coff_l;
  int* cell_of = shortlist ? c->cell_of_s : c->cell_of_l;
  int* cnt = shortlist ? c->cnt_s : c->cnt_l;
  int* start = shortlist ? c->start_s : c->start_l;
  int* cat = shortlist ? c->atoms_s : c->atoms_l;
  double rc = shortlist ? c->rc_short : c->rc_long;
  double rr = (1.0 + 0.44) * rc;  // NeighbourList.cxx:62
  double range2 = rr * rr;
  int n = c->n;
  Atoms& A = c->A;
  int L = shortlist ? 0 : 1;
  int nsys = c->nsys;
  launch(c, "k_cell_clear",
         [&] { k_cell_clear<<<nsys, 256, 0, c->stream>>>(flag, c->sflags, L, coff, cnt); });
  launch(c, "k_cell_assign", [&] {
    k_cell_assign<<<blocks(n), 256, 0, c->stream>>>(flag, c->sflags, L, n, A.p, g, coff, A.sys, cell_of, cnt);
  });
  launch(c, "k_cell_scan", [&] {
    k_cell_scan<<<nsys, 1024, 0, c->stream>>>(flag, c->sflags, L, coff, A.sys_off, cnt, start);
  });
  launch(c, "k_cell_fill", [&] {
    k_cell_fill<<<blocks(n), 256, 0, c->stream>>>(flag, c->sflags, L, n, A.sys, cell_of, start, cnt, cat);
  });
  // systems whose atoms carry their own PBC vectors get their lists from
  // the APBC instantiation (all pairs, each atom's own depos)
  const bool apbc = c->box.apbc != nullptr;
  if (shortlist) {
    launch(c, "k_nl_build_short", [&] {
      auto go = [&](auto ap) {
        k_nl_build<64, true, decltype(ap)::value><<<blocks(n, 4, 8192), 256, 0, c->stream>>>(
            flag, c->sflags, 0, n, A.p, c->handled_s, c->box, g, coff, range2, cell_of, start, cat, A.nls_cnt,
            A.nls, A.disp_s, c->derr, nullptr, A.x, A.y, A.z);
      };
      go(std::false_type());
      if (apbc) go(std::true_type());
    });
  } else {
    launch(c, "k_nl_build_long", [&] {
      auto go = [&](auto sorted, auto ap) {
        k_nl_build<512, decltype(sorted)::value, decltype(ap)::value><<<blocks(n, 4, 8192), 256, 0, c->stream>>>(
            flag, c->sflags, 1, n, A.p, c->handled_l, c->box, g, coff, range2, cell_of, start, cat, A.nll_cnt,
            A.nll, A.disp_l, c->derr, A.nll_split, A.x, A.y, A.z);
      };
      if (c->mixed) {  // rows split at j > i only (no sort)
        go(std::false_type(), std::false_type());
        if (apbc) go(std::false_type(), std::true_type());
      } else {
        go(std::true_type(), std::false_type());
        if (apbc) go(std::true_type(), std::true_type());
      }
    });
    launch(c, "k_long_nbflag", [&] {
      k_long_nbflag<<<blocks(n, 16), 256, 0, c->stream>>>(flag, c->sflags, A, c->box, c->lmem);
    });
  }
}

// The differentiable parameters theta (kNumTheta values, CGS) inside the FP64
// tables: for p = C-C, C-H, H-H, REBO's {Q, A, alpha, B1-3, beta1-3} at 9p
// (REBO.cxx:856-941) and LJ {epsilon, sigma} at 27 + 2p (AIREBO_LJ.cxx:229-247).
// write = false reads them out.  sigma also sets the S_r range [sigma,
// 2^(1/6) sigma] (AIREBO_LJ.cxx:239-247).
void theta_io(Tables& T, double* th, bool write) {
  const int pa[3] = {T_C, T_C, T_H}, pb[3] = {T_C, T_H, T_H};
  for (int p = 0; p < 3; p++) {
    int a = pa[p], b = pb[p];
    double* v[9] = {&T.rebo.Q[a][b], &T.rebo.A[a][b], &T.rebo.alpha[a][b], &T.rebo.B[0][a][b], &T.rebo.B[1][a][b],
                    &T.rebo.B[2][a][b], &T.rebo.beta[0][a][b], &T.rebo.beta[1][a][b], &T.rebo.beta[2][a][b]};
    double* w[9] = {&T.rebo.Q[b][a], &T.rebo.A[b][a], &T.rebo.alpha[b][a], &T.rebo.B[0][b][a], &T.rebo.B[1][b][a],
                    &T.rebo.B[2][b][a], &T.rebo.beta[0][b][a], &T.rebo.beta[1][b][a], &T.rebo.beta[2][b][a]};
    for (int k = 0; k < 9; k++) {
      if (write) *v[k] = *w[k] = th[9 * p + k];
      else th[9 * p + k] = *v[k];
    }
    if (write) {
      T.lj.zeta[a][b] = T.lj.zeta[b][a] = th[27 + 2 * p];
      T.lj.sigma[a][b] = T.lj.sigma[b][a] = th[27 + 2 * p + 1];
      for (int q = 0; q < 2; q++) {
        int x = q ? b : a, y = q ? a : b;
        T.lj.RLJ[x][y][0] = T.lj.sigma[x][y];
        T.lj.RLJ[x][y][1] = T.lj.sigma[x][y] * std::pow(2.0, 1.0 / 6.0);
      }
    } else {
      th[27 + 2 * p] = T.lj.zeta[a][b];
      th[27 + 2 * p + 1] = T.lj.sigma[a][b];
    }
  }
}

// Clears one list's rebuild flags (all systems).
__global__ void k_clear_list(int nsys, int list, int* flags, int* sflags) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) flags[list] = 0;
  if (s < nsys) sflags[2 * s + list] = 0;
}

__global__ void k_fill(double* p, int n, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Clears the per-call tallies (LJ pairs, quads) of every slot; blockDim kTallySlots.
__global__ void k_zero_counts(long long* tally) {
  for (int k : {0, 1, 3, 4, 5}) tally[threadIdx.x * kTallyStride + k] = 0;  // [2] counts across calls
}

// The bond-order chain members, in the geometry type R of the context.  VIR:
// the kernels' virial variants (one launch shape each), which also add the
// per-system virial into c->vir.
template <class R, bool TH, bool VIR>
void force_terms(mdg_ctx* c, const TablesT<R>* T) {
  static_assert(!(TH && VIR), "parameter-gradient calls do not reduce the virial");
  int n = c->n;
  Atoms& A = c->A;
  double* dth = TH ? c->dtheta : nullptr;
  long long* vir = VIR ? c->virfix : nullptr;
  if (c->have_short) {
    launch(c, "k_coord", [&] { k_coord<R><<<blocks(n, 16), 256, 0, c->stream>>>(A, T, c->box, c->derr); });
    launch(c, "k_bond_enum", [&] {
      k_bond_enum<<<blocks(n), 256, 0, c->stream>>>(A, c->bonds, c->nbonds, c->bond_cap, c->derr, c->tally + 3);
    });
  }
  const Tune& tu = c->tune;
  int bgrid = tu.grid_bonds ? tu.grid_bonds : blocks(2LL * n, 256, 8192);
  int tgrid = tu.grid_tors ? tu.grid_tors : blocks(2LL * n, 256, 8192);
  int sgrid = tu.grid_bstar ? tu.grid_bstar : blocks(8LL * n, 256, 16384);
  // The bond-centred kernels (REBO bonds, Brenner, torsion) only read the
  // bond lists and add into grad, so they run on the side stream beside the
  // LJ chain (pair rows -> LJ -> b*); an event joins them back.  Profiled
  // calls stay on one stream so that every kernel's own time is measured.
  const bool fork = !c->profiling && c->side && c->slot_lj >= 0 &&
                    (c->slot_rebo >= 0 || c->slot_bren >= 0 || c->slot_tors >= 0);
  hipStream_t sb = fork ? c->side : c->stream;
  if (fork) {
    hipEventRecord(c->ev_fork, c->stream);
    hipStreamWaitEvent(c->side, c->ev_fork, 0);
  }
  if (c->slot_rebo >= 0) {
    launch(c, "k_rebo_bonds", [&] {
      auto go = [&](auto occ) {
        k_rebo_bonds<R, TH, decltype(occ)::value, VIR><<<bgrid, 256, 0, sb>>>(
            A, T, c->dT, c->box, c->bonds, c->nbonds, A.grad, c->eatom + (size_t)c->slot_rebo * n, c->derr, dth,
            vir);
      };
      if constexpr (!TH && !VIR) {
        if (tu.occ_bonds >= 11) {  // LDS-staged shells, one wavefront per tile of 64 bonds
          auto sh = [&](auto occ) {
            k_rebo_shell<R, decltype(occ)::value><<<bgrid, 256, 0, sb>>>(A, T, c->dT, c->box, c->bonds, c->nbonds,
                                                                         A.grad,
                                                                         c->eatom + (size_t)c->slot_rebo * n, c->derr);
          };
          if constexpr (std::is_same<R, float>::value) {
            if (tu.occ_bonds == 12) return sh(std::integral_constant<int, 2>());
          }
          return sh(std::integral_constant<int, 1>());
        }
        if (tu.occ_bonds == 2) return go(std::integral_constant<int, 2>());
        if constexpr (std::is_same<R, float>::value) {  // 3-4 waves/SIMD: FP32 form only
          if (tu.occ_bonds == 3) return go(std::integral_constant<int, 3>());
          if (tu.occ_bonds == 4) return go(std::integral_constant<int, 4>());
        }
      }
      go(std::integral_constant<int, 1>());
    });
  }
  if (c->slot_bren >= 0) {
    launch(c, "k_brenner_bonds", [&] {
      k_brenner_bonds<VIR><<<bgrid, 256, 0, sb>>>(A, c->dT, c->box, c->bonds, c->nbonds, A.grad,
                                                  c->eatom + (size_t)c->slot_bren * n, c->derr, vir);
    });
  }
  if (c->slot_tors >= 0) {
    launch(c, "k_etors", [&] {
      k_etors<R, VIR><<<tgrid, 256, 0, sb>>>(A, T, c->box, c->bonds, c->nbonds, c->cfg.etors_compat, A.grad,
                                             c->eatom + (size_t)c->slot_tors * n, c->tally + 1, vir);
    });
  }
  if (c->slot_lj >= 0) {
    launch(c, "k_pair_rows", [&] {
      k_pair_rows<R><<<blocks(n, 16, 16384), 256, 0, c->stream>>>(A, T, c->box, c->flags + 2, c->pflags,
                                                                 c->pair_margin, c->tally + 2);
    });
    launch(c, "k_lj_wave", [&] {
      auto go = [&](auto occ, auto lw) {
        constexpr int LW = decltype(lw)::value;
        int grid = tu.grid_lj ? tu.grid_lj : blocks(n, 4 * (64 / LW), 65536);
        k_lj_wave<R, TH, decltype(occ)::value, LW, VIR><<<grid, 256, 0, c->stream>>>(
            A, T, c->box, A.grad, c->eatom + (size_t)c->slot_lj * n, c->queue, c->qcount, c->qsub,
            c->tally + 0, c->derr, dth, vir);
      };
      using I1 = std::integral_constant<int, 1>;
      if constexpr (TH || VIR) {
        go(I1(), std::integral_constant<int, 16>());
      } else {
        switch (tu.lj_shape) {
          case 0: go(I1(), std::integral_constant<int, 64>()); break;
          case 1: go(std::integral_constant<int, 3>(), std::integral_constant<int, 64>()); break;
          case 2: go(I1(), std::integral_constant<int, 32>()); break;
          case 4:  // 16 lanes at 4 or 3 waves/SIMD (FP32 form only: the FP64 form's LDS allows 2)
          case 5:
            if constexpr (std::is_same<R, float>::value) {
              if (tu.lj_shape == 4) go(std::integral_constant<int, 4>(), std::integral_constant<int, 16>());
              else go(std::integral_constant<int, 3>(), std::integral_constant<int, 16>());
              break;
            }
            [[fallthrough]];
          default: go(I1(), std::integral_constant<int, 16>()); break;
        }
      }
    });
    launch(c, "k_bstar", [&] {
      double* eb = c->eatom + (size_t)c->slot_lj * n;
      if constexpr (TH || VIR) {  // one pass
        k_bstar<R, TH, 1, 0, VIR><<<sgrid, 256, 0, c->stream>>>(A, T, c->box, c->queue, c->qcount, c->qsub, A.grad,
                                                                 eb, c->derr, dth, vir, c->tally + 4);
      } else {
        // the bound pass at the tuned occupancy, then the full pass over the
        // pairs whose bound failed
        auto go = [&](auto occ) {
          k_bstar<R, false, decltype(occ)::value, 1, false><<<sgrid, 256, 0, c->stream>>>(
              A, T, c->box, c->queue, c->qcount, c->qsub, A.grad, eb, c->derr, dth, nullptr, c->tally + 4);
        };
        int occ = tu.occ_bstar;
        if constexpr (std::is_same<R, float>::value) {  // 3-4 waves/SIMD: FP32 form only
          if (occ == 3) go(std::integral_constant<int, 3>());
          if (occ == 4) go(std::integral_constant<int, 4>());
        } else {
          if (occ > 2) occ = 1;
        }
        if (occ == 2) go(std::integral_constant<int, 2>());
        if (occ == 1) go(std::integral_constant<int, 1>());
        k_bstar<R, false, 1, 2, false><<<sgrid, 256, 0, c->stream>>>(A, T, c->box, c->queue, c->qcount, c->qsub,
                                                                      A.grad, eb, c->derr, dth, nullptr, c->tally + 4);
      }
    });
  }
  if (fork) {
    hipEventRecord(c->ev_join, c->side);
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
  }
}

static_assert(sizeof(mdg_frame) == sizeof(TrajFrame), "mdg_frame layout");
static_assert(MDG_ERR_SHORT_NL_OVERFLOW == ERR_SHORT_NL_OVERFLOW && MDG_ERR_LONG_NL_OVERFLOW == ERR_LONG_NL_OVERFLOW &&
                  MDG_ERR_BOND_OVERFLOW == ERR_BOND_OVERFLOW && MDG_ERR_COS_RANGE == ERR_COS_RANGE &&
                  MDG_ERR_NEG_COORD == ERR_NEG_COORD && MDG_ERR_QUEUE_OVERFLOW == ERR_QUEUE_OVERFLOW &&
                  MDG_ERR_SPLINE_DOMAIN == ERR_SPLINE_DOMAIN && MDG_ERR_EAM_ROW_OVERFLOW == ERR_EAM_ROW_OVERFLOW &&
                  MDG_ERR_PATH_OVERFLOW == ERR_PATH_OVERFLOW && MDG_ERR_TYPE_RANGE == ERR_TYPE_RANGE,
              "public error bits");
static_assert(sizeof(mdg_atom_record) == sizeof(TrajRec), "mdg_atom_record layout");
constexpr int kTrajChunk = 256;  // steps between drains of the trajectory buffers

// Appends the frames and records captured since the chunk started to the
// host lists (frames' first indexes re-based on the host list).
int drain_trajectory(mdg_ctx* c) {
  unsigned long long cnt[3];
  HIPCHK(c, hipMemcpy(cnt, c->tb.cnt, sizeof(cnt), hipMemcpyDeviceToHost));
  if (cnt[2]) {
    c->err = "trajectory output buffer overflow";
    return 1;
  }
  size_t nf = cnt[0], nr = cnt[1], f0 = c->frames_h.size(), r0 = c->recs_h.size();
  c->frames_h.resize(f0 + nf);
  c->recs_h.resize(r0 + nr);
  if (nf) HIPCHK(c, hipMemcpy(&c->frames_h[f0], c->tb.frm, nf * sizeof(mdg_frame), hipMemcpyDeviceToHost));
  if (nr) HIPCHK(c, hipMemcpy(&c->recs_h[r0], c->tb.rec, nr * sizeof(mdg_atom_record), hipMemcpyDeviceToHost));
  for (size_t k = f0; k < f0 + nf; k++) c->frames_h[k].first += (int64_t)r0;
  return 0;
}

// The skin test of every system (both stages).
void sys_disp(mdg_ctx* c, double lim_s, double lim_l) {
  Atoms& A = c->A;
  k_sys_disp_part<<<c->nred, 256, 0, c->stream>>>(c->rb, A.sys_off, A.disp_s, A.disp_l, A.disp_p, A.disp_e,
                                                 c->red_part);
  k_sys_disp_merge<<<c->nsys, 64, 0, c->stream>>>(c->rb_off, c->red_part, lim_s, lim_l, c->pair_margin,
                                                  A.disp_p ? 1 : 0, c->metal_skin, A.disp_e ? 1 : 0, c->sflags,
                                                  c->pflags, c->eflags, c->flags, c->counters);
}

// One force call (FProxy::operator(), FProxy.cxx:45-61): lists, coordination,
// then every chain member adds into grad and its energy slot.
void force_call(mdg_ctx* c) {
  int n = c->n;
  Atoms& A = c->A;
  if (c->have_short) nl_build(c, true);
  if (c->have_long) nl_build(c, false);
  size_t nt = c->kinds.size();
  launch(c, "k_zero", [&] {
    if (A.gfix) {  // all-zero bits are integer zero too
      k_zero<<<blocks(6LL * n, 256, 4096), 256, 0, c->stream>>>(reinterpret_cast<double*>(A.gfix), (size_t)6 * n);
      k_zero<<<blocks(2LL * nt * n, 256, 4096), 256, 0, c->stream>>>(reinterpret_cast<double*>(A.efix),
                                                                     2 * nt * n);
    } else {
      k_zero<<<blocks(3LL * n, 256, 4096), 256, 0, c->stream>>>(A.grad, (size_t)3 * n);
    }
    k_zero<<<blocks((long long)nt * n, 256, 4096), 256, 0, c->stream>>>(c->eatom, nt * n);
  });
  k_zero_i<<<1, 64, 0, c->stream>>>(c->nbonds, 1);
  k_zero_i<<<kQueues * kQStride / 256, 256, 0, c->stream>>>(c->qcount, kQueues * kQStride);
  k_zero_counts<<<1, kTallySlots, 0, c->stream>>>(c->tally);
  const bool vir = c->virial && !c->theta_mode;
  long long* vw = vir ? c->virfix : nullptr;
  if (vir) k_zero<<<1, 64, 0, c->stream>>>(reinterpret_cast<double*>(c->virfix), 12 * (size_t)c->nsys);  // all-zero bits
  if (c->mixed) {
    if (vir) force_terms<float, false, true>(c, c->dTf);
    else force_terms<float, false, false>(c, c->dTf);
  } else if (c->theta_mode) {
    force_terms<double, true, false>(c, c->dT);
  } else {
    if (vir) force_terms<double, false, true>(c, c->dT);
    else force_terms<double, false, false>(c, c->dT);
  }
  if (c->slot_flj >= 0) {
    launch(c, "k_flj", [&] {
      double* ea = c->eatom + (size_t)c->slot_flj * n;
      if (vir) k_flj<true><<<blocks(n), 256, 0, c->stream>>>(A, c->box, c->dFLJ, A.grad, ea, vw);
      else k_flj<false><<<blocks(n), 256, 0, c->stream>>>(A, c->box, c->dFLJ, A.grad, ea, nullptr);
    });
  }
  if (A.ml) {
    launch(c, "k_metal_rows", [&] {
      k_metal_rows<<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->flags + 3, c->eflags, c->metal_rcand);
    });
  }
  if (c->slot_eam >= 0) {
    double* ea = c->eatom + (size_t)c->slot_eam * n;
    const auto& E = c->emb[0];
    launch(c, "k_eam_density", [&] {
      k_eam_density<AcklandModel><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dEAM, E.e, E.fe, ea, E.row,
                                                                        E.cnt, c->derr);
    });
    launch(c, "k_eam_force", [&] {
      if (vir)
        k_eam_force<AcklandModel, true><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dEAM, E.e, E.fe, E.row,
                                                                              E.cnt, A.grad, ea, vw);
      else
        k_eam_force<AcklandModel, false><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dEAM, E.e, E.fe, E.row,
                                                                               E.cnt, A.grad, ea, nullptr);
    });
  }
  if (c->slot_fbm >= 0) {
    launch(c, "k_fbm", [&] {
      double* ea = c->eatom + (size_t)c->slot_fbm * n;
      if (vir) k_fbm<true><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dFBM, A.grad, ea, vw);
      else k_fbm<false><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dFBM, A.grad, ea, nullptr);
    });
  }
  if (c->slot_tb >= 0) {
    double* ea = c->eatom + (size_t)c->slot_tb * n;
    const auto& E = c->emb[1];
    launch(c, "k_tb_density", [&] {
      k_eam_density<TBModel><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dTB, E.e, E.fe, ea, E.row, E.cnt,
                                                                   c->derr);
    });
    launch(c, "k_tb_force", [&] {
      if (vir)
        k_eam_force<TBModel, true><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dTB, E.e, E.fe, E.row, E.cnt,
                                                                         A.grad, ea, vw);
      else
        k_eam_force<TBModel, false><<<blocks(n, 16), 256, 0, c->stream>>>(A, c->box, c->dTB, E.e, E.fe, E.row,
                                                                          E.cnt, A.grad, ea, nullptr);
    });
  }
  if (c->slot_fbzl >= 0) {
    launch(c, "k_fbzl", [&] {
      double* ea = c->eatom + (size_t)c->slot_fbzl * n;
      if (c->nions == 0) return;
      const int g = blocks(64LL * c->nions);  // a wave per ion
      if (vir)
        k_fbzl<true><<<g, 256, 0, c->stream>>>(A, c->box, c->ions, c->nions, c->cfg.fbzl_r0, c->cfg.fbzl_r1, A.grad,
                                               ea, vw);
      else
        k_fbzl<false><<<g, 256, 0, c->stream>>>(A, c->box, c->ions, c->nions, c->cfg.fbzl_r0, c->cfg.fbzl_r1,
                                                A.grad, ea, nullptr);
    });
  }
  if (A.gfix) {  // the fixed-point sums into grad and the atomically summed energy slots
    unsigned fixmask = 0;
    for (int t = 0; t < (int)nt; t++)
      if (t != c->slot_eam && t != c->slot_fbm && t != c->slot_tb) fixmask |= 1u << t;  // those store directly
    launch(c, "k_fix_join", [&] {
      k_fix_join<<<blocks((3LL + nt) * n, 256, 8192), 256, 0, c->stream>>>(A, (int)nt, fixmask, c->eatom);
    });
  }
  if (vir) k_vir_join<<<blocks(6LL * c->nsys), 256, 0, c->stream>>>(c->nsys, c->virfix, c->vir);
  k_clear_flags<<<blocks(2LL * c->nsys), 256, 0, c->stream>>>(c->nsys, c->flags, c->sflags, c->pflags, c->eflags);
}

// A copy of a host object into a new device buffer of the context.
template <class T>
int upload(mdg_ctx* c, T*& d, const T& h) {
  if (dalloc(c, d, 1)) return 1;
  HIPCHK(c, hipMemcpy(d, &h, sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// mdg_create's work on a constructed context: the potential chain in the
// order of applications/common.h:46-68, the parameter tables, streams and
// the device counters.  With c->arena set (mdg_energy_grad_noalloc) every
// device buffer comes from that memory, and no stream or event is created:
// the context works on the null stream without the side stream.
int ctx_init(mdg_ctx* c, const mdg_config* cfg) {
  if (!cfg) {
    c->err = "null config";
    return 1;
  }
  int m = cfg->pot_mask;
  if ((m & MDG_POT_AIREBO) && !(m & MDG_POT_REBO)) {
    c->err = "AIREBO-LJ needs REBO in the chain (AIREBO_LJ.hpp:63)";
    return 1;
  }
  if ((m & MDG_POT_BRENNER) && (m & (MDG_POT_REBO | MDG_POT_AIREBO))) {
    c->err = "Brenner replaces REBO: select one of them";
    return 1;
  }
  if (!(m & (MDG_POT_REBO | MDG_POT_BRENNER | MDG_POT_ETORS | MDG_POT_FBZL | MDG_POT_FLJ | MDG_POT_ACKLAND |
             MDG_POT_FBM | MDG_POT_TB))) {
    c->err = "no potentials selected";
    return 1;
  }
  if (cfg->precision != MDG_PREC_FP64 && cfg->precision != MDG_PREC_MIXED) {
    c->err = "precision must be MDG_PREC_FP64 or MDG_PREC_MIXED";
    return 1;
  }
  if (cfg->precision == MDG_PREC_MIXED && (m & MDG_POT_BRENNER)) {
    c->err = "the mixed-precision form covers REBO/AIREBO/ETors/FBZL; run Brenner in FP64";
    return 1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    c->err = "no HIP device";
    return 1;
  }
  c->cfg = *cfg;
  c->use_graphs = std::getenv("MDG_NO_GRAPH") == nullptr;
  if (c->cfg.max_short_nb <= 0) c->cfg.max_short_nb = 64;
  if (c->cfg.max_long_nb <= 0) c->cfg.max_long_nb = 512;
  if (c->cfg.max_short_nb != 64 || c->cfg.max_long_nb != 512) {
    c->err = "list capacities are fixed at 64 (short) and 512 (long) in this library";
    return 1;
  }
  if (c->cfg.fbzl_r1 <= 0) {
    c->cfg.fbzl_r0 = 5.0 * kAo;
    c->cfg.fbzl_r1 = 5.5 * kAo;
  }
  // chain order of applications/common.h:46-68
  if (m & MDG_POT_FBZL) { c->slot_fbzl = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_FBZL); }
  if (m & MDG_POT_FLJ) { c->slot_flj = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_FLJ); }
  if (m & MDG_POT_BRENNER) { c->slot_bren = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_BRENNER); }
  if (m & MDG_POT_REBO) { c->slot_rebo = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_REBO); }
  if (m & MDG_POT_AIREBO) { c->slot_lj = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_AIREBO); }
  if (m & MDG_POT_ETORS) { c->slot_tors = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_ETORS); }
  if (m & MDG_POT_ACKLAND) { c->slot_eam = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_ACKLAND); }
  // FBM and TightBinding are not registered by common.h (:69-72); selected, they follow the chain
  if (m & MDG_POT_FBM) { c->slot_fbm = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_FBM); }
  if (m & MDG_POT_TB) { c->slot_tb = (int)c->kinds.size(); c->kinds.push_back(MDG_POT_TB); }
  c->have_short = (m & (MDG_POT_REBO | MDG_POT_BRENNER | MDG_POT_ETORS)) != 0;
  c->have_long = (m & (MDG_POT_AIREBO | MDG_POT_FBZL | MDG_POT_FLJ | MDG_POT_ACKLAND | MDG_POT_FBM | MDG_POT_TB)) != 0;
  c->rc_short = 2.0 * kAo;  // REBO/Brenner/ETors getRcutoff (REBO.hpp:130)
  c->rc_long = 0;
  if (m & MDG_POT_AIREBO) c->rc_long = 5.5 * kAo;  // AIREBO_LJ.hpp:81
  if (m & MDG_POT_FBZL) c->rc_long = std::max(c->rc_long, c->cfg.fbzl_r1);
  if (m & MDG_POT_FLJ) c->rc_long = std::max(c->rc_long, 5.5 * kAo);  // common.h:49
  EAMTable et;
  build_eam_table(et);
  // Ackland::getRcutoff (Ackland.hpp:58): the largest like-pair first knot
  if (m & MDG_POT_ACKLAND) c->rc_long = std::max(c->rc_long, std::max(et.rk[0][0][1], std::max(et.rk[1][1][1], et.rk[2][2][1])));
  TBTable tt;
  build_tb_table(tt);
  FBMTable ft2;
  build_fbm_table(ft2);
  if (m & MDG_POT_FBM) c->rc_long = std::max(c->rc_long, ft2.R1);  // FPairwise::getRcutoff
  if (m & MDG_POT_TB) c->rc_long = std::max(c->rc_long, tt.R1);    // TightBinding.hpp:55
  c->metal_rcand = std::max((m & MDG_POT_ACKLAND) ? et.rmax : 0.0, (m & MDG_POT_TB) ? tt.R1 : 0.0) + c->metal_skin;
  {  // handled elements (H, C, Ar, Xe, Cu, Ag, Au bits) and list radius of each long-range member
    auto add = [&](unsigned h, double rc) {
      c->lmem.handled[c->lmem.m] = h;
      c->lmem.r2[c->lmem.m] = (1.44 * rc) * (1.44 * rc);
      c->lmem.m++;
    };
    if (m & MDG_POT_FBZL) add(0x7Fu, c->cfg.fbzl_r1);  // FBZL.cxx:36-58
    if (m & MDG_POT_AIREBO) add(0x03u, 5.5 * kAo);     // AIREBO_LJ: H, C
    if (m & MDG_POT_FLJ) add(0x53u, 5.5 * kAo);        // FLJ.cxx:38-52: H, C, Cu, Au
    if (m & MDG_POT_ACKLAND)                           // Ackland.cxx:236-262: Cu, Ag, Au
      add(0x70u, std::max(et.rk[0][0][1], std::max(et.rk[1][1][1], et.rk[2][2][1])));
    if (m & MDG_POT_FBM) add(0x14u, ft2.R1);           // FBM.cxx:37-38: Cu, Ar
    if (m & MDG_POT_TB) add(0x10u, tt.R1);             // TightBinding.cxx:163: Cu
  }
  if (!c->arena) {
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
      c->err = "hipStreamCreate failed";
      return 1;
    }
    c->stream = c->own;
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
      c->err = "hipStreamCreate/hipEventCreate failed";
      return 1;
    }
  }
  Tables h;
  build_tables(h, cfg->brenner_paramset);
  c->host_T = h;
  if (upload(c, c->dT, h)) return 1;
  c->mixed = cfg->precision == MDG_PREC_MIXED;
  if (c->mixed) {
    TablesT<float> hf;
    convert_tables(h, hf);
    if (upload(c, c->dTf, hf)) return 1;
  }
  if (c->slot_flj >= 0) {
    FLJTable ft;
    build_flj_table(ft);
    if (upload(c, c->dFLJ, ft)) return 1;
  }
  if ((c->slot_eam >= 0 && upload(c, c->dEAM, et)) || (c->slot_tb >= 0 && upload(c, c->dTB, tt)) ||
      (c->slot_fbm >= 0 && upload(c, c->dFBM, ft2)))
    return 1;
  if (dalloc(c, c->flags, 4) || dalloc(c, c->counters, 8) || dalloc(c, c->tally, kTallySlots * kTallyStride) ||
      dalloc(c, c->derr, 1) || dalloc(c, c->nbonds, 1) || dalloc(c, c->qcount, kQueues * kQStride))
    return 1;
  HIPCHK(c, hipMemset(c->counters, 0, 8 * sizeof(long long)));
  HIPCHK(c, hipMemset(c->tally, 0, kTallySlots * kTallyStride * sizeof(long long)));
  HIPCHK(c, hipMemset(c->derr, 0, sizeof(unsigned)));
  return 0;
}
}  // namespace

extern "C" {

const char* mdg_last_error(const mdg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int mdg_create(const mdg_config* cfg, mdg_ctx** out) {
  *out = nullptr;
  mdg_ctx* c = new mdg_ctx();
  if (ctx_init(c, cfg)) {
    g_create_err = c->err;
    delete c;
    return 1;
  }
  *out = c;
  return 0;
}

void mdg_destroy(mdg_ctx* ctx) { delete ctx; }

int mdg_num_terms(const mdg_ctx* ctx) { return (int)ctx->kinds.size(); }
int mdg_term_kind(const mdg_ctx* ctx, int i) {
  return (i >= 0 && i < (int)ctx->kinds.size()) ? ctx->kinds[i] : -1;
}

int mdg_set_stream(mdg_ctx* c, void* s) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stream = s ? static_cast<hipStream_t>(s) : c->own;
  return 0;
}

int mdg_set_systems(mdg_ctx* c, int nsys, const int32_t* counts, const int32_t* elem, const double* xyz,
                    const double* vel, const double* pbc, const uint8_t* fixed) {
  if (nsys <= 0) {
    c->err = "no systems";
    return 1;
  }
  if (c->n != 0) {
    c->err = "mdg_set_system(s): a context holds one system or one batch; create a new context";
    return 1;
  }
  std::vector<int> off(nsys + 1, 0);
  for (int s = 0; s < nsys; s++) {
    if (counts[s] <= 0) {
      c->err = "empty system (AtomsContainer.cxx:108)";
      return 1;
    }
    if ((long long)off[s] + counts[s] > (1LL << 30)) {
      c->err = "too many atoms in one context";
      return 1;
    }
    off[s + 1] = off[s] + counts[s];
  }
  int n = off[nsys];
  double rcmax = std::max(c->rc_short, c->rc_long);  // checkMIC(2 Rc), SimLoop.cxx:239-245
  for (int s = 0; s < nsys; s++)
    for (int d = 0; d < 3; d++)
      if (pbc[3 * s + d] <= rcmax * 2.0) {
        c->err = "Rcutoff is too large for given PBC ! (SimLoop.cxx:239-245)";
        if (nsys > 1) c->err += " [system " + std::to_string(s) + "]";
        return 1;
      }
  std::vector<double> p(xyz, xyz + 3 * (size_t)n);
  std::vector<int> pc(3 * (size_t)n, 0), ty(n), el(n), sy(n);
  std::vector<double> ms(n), zs(n);
  std::vector<unsigned char> fx(n, 0), hs(n), hl(n);
  for (int s = 0; s < nsys; s++)
    for (int i = off[s]; i < off[s + 1]; i++) {
      sy[i] = s;
      el[i] = elem[i];
      ty[i] = type_of(elem[i]);
      ms[i] = mass_of(elem[i]);
      zs[i] = charge_of(elem[i]);
      if (ms[i] <= 0) {
        c->err = "unknown element id " + std::to_string(elem[i]);
        return 1;
      }
      fx[i] = fixed ? (fixed[i] ? 1 : 0) : 0;
      hs[i] = (ty[i] <= T_H) ? 1 : 0;
      hl[i] = 1;
      for (int d = 0; d < 3; d++) {  // Atom::applyPBC (Atom.cxx:93-125)
        double L = pbc[3 * s + d];
        if (L == kNoPbc) continue;
        double& q = p[3 * i + d];
        int& k = pc[3 * i + d];
        int guard = 0;
        while (q < 0 && guard++ < 1000000) { q += L; --k; }
        while (q >= L && guard++ < 1000000) { q -= L; ++k; }
      }
    }
  c->n = n;
  c->nsys = nsys;
  c->off_h = off;
  c->failed.assign(nsys, 0u);
  {  // error words: [0] global resources, [1 + s] system s
    dfree(c, c->derr);
    if (dalloc(c, c->derr, 1 + (size_t)nsys)) return 1;
    HIPCHK(c, hipMemset(c->derr, 0, (1 + (size_t)nsys) * sizeof(unsigned)));
  }
  c->Lh.assign(pbc, pbc + 3 * (size_t)nsys);
  c->dt_host.assign(nsys, 1e-20);  // SimLoop.cxx:50
  Atoms& A = c->A;
  A.n = n;
  A.nsys = nsys;
  A.cap_s = c->cfg.max_short_nb;
  A.cap_l = c->cfg.max_long_nb;
  A.maxb = 16;
  int r = 0;
  double* Ld = nullptr;
  int* sysd = nullptr;
  r |= dalloc(c, Ld, 3 * (size_t)nsys) | dalloc(c, sysd, n) | dalloc(c, A.sys_off, nsys + 1);
  A.sys = sysd;
  c->box.L = Ld;
  c->box.sys = sysd;
  r |= dalloc(c, c->sflags, 2 * (size_t)nsys) | dalloc(c, c->pflags, nsys) | dalloc(c, c->dtp, nsys) | dalloc(c, c->eacc, 8 * (size_t)nsys);
  r |= dalloc(c, c->eatom, c->kinds.size() * (size_t)n);
  A.eatom0 = c->eatom;
  A.gfix = nullptr;
  A.efix = nullptr;
  if (!c->mixed)  // the FP64 form accumulates in fixed point (deterministic)
    r |= dalloc(c, A.gfix, 6 * (size_t)n) | dalloc(c, A.efix, 2 * c->kinds.size() * (size_t)n);
  r |= dalloc(c, A.x, n) | dalloc(c, A.y, n) | dalloc(c, A.z, n);
  r |= dalloc(c, A.vx, n) | dalloc(c, A.vy, n) | dalloc(c, A.vz, n);
  r |= dalloc(c, A.ax, n) | dalloc(c, A.ay, n) | dalloc(c, A.az, n);
  r |= dalloc(c, A.grad, 3 * (size_t)n) | dalloc(c, A.mass, n) | dalloc(c, A.Z, n);
  r |= dalloc(c, A.an_nt, 3 * (size_t)n);
  A.bath = nullptr;  // a new set of systems starts without a bath
  A.tb_apply = nullptr;
  A.tb_T = nullptr;
  A.tb_dE = nullptr;
  c->bath_d = nullptr;
  c->tb_apply_d = nullptr;
  c->ebath = nullptr;
  c->bath_on = false;
  c->bath_h.assign(nsys, Bath{TB_NONE, 0.0, 1.0e13, 1000000.0 * kAo, 0.0, -5.0 * kAo, 0.0, 0.0, 0.0,
                              1000000.0 * kAo});
  r |= dalloc(c, A.type, n) | dalloc(c, A.elem, n) | dalloc(c, A.fixed, n);
  {  // the ZBL term's centres
    std::vector<int> ions;
    for (int i = 0; i < n; i++)
      if (el[i] == 40 || el[i] == 131) ions.push_back(i);
    c->nions = (int)ions.size();
    c->ions = nullptr;
    if (c->nions) {
      r |= dalloc(c, c->ions, ions.size());
      if (!r) HIPCHK(c, hipMemcpy(c->ions, ions.data(), ions.size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  r |= dalloc(c, A.pcx, n) | dalloc(c, A.pcy, n) | dalloc(c, A.pcz, n);
  r |= dalloc(c, A.disp_s, 3 * (size_t)n) | dalloc(c, A.disp_l, 3 * (size_t)n);
  r |= dalloc(c, A.nls_cnt, n) | dalloc(c, A.nll_cnt, n);
  if (c->have_short) r |= dalloc(c, A.nls, (size_t)n * A.cap_s);
  A.nbl = nullptr;
  if (c->have_long) r |= dalloc(c, A.nll, (size_t)n * A.cap_l) | dalloc(c, A.nll_split, n) | dalloc(c, A.nbl, n);
  r |= dalloc(c, A.bcnt, n) | dalloc(c, A.bj, (size_t)n * A.maxb);
  r |= dalloc(c, A.p, n);
  if (!c->mixed) r |= dalloc(c, A.b64, (size_t)n * A.maxb);
  r |= dalloc(c, A.Nt, n) | dalloc(c, A.NH, n) | dalloc(c, A.NC, n);
  A.pmax = nullptr;
  if (c->slot_lj >= 0) r |= dalloc(c, A.pmax, 2 * (size_t)n);
  if (c->mixed) {
    r |= dalloc(c, A.b32, (size_t)n * A.maxb);
  }
  c->bond_cap = n * A.maxb / 2 + 64;
  A.bond_cap = c->bond_cap;
  r |= dalloc(c, c->bonds, c->bond_cap);
  c->qsub = (32 * n + kQueues - 1) / kQueues + 4096;  // >= 16 atoms x 128 pairs for any one workgroup pass
  if (c->slot_lj >= 0) r |= dalloc(c, c->queue, (size_t)c->qsub * kQueues);
  A.cap_p = 128;
  A.pl = nullptr;
  A.pl_cnt = nullptr;
  A.disp_p = nullptr;
  if (c->slot_lj >= 0)
    r |= dalloc(c, A.pl, (size_t)n * A.cap_p) | dalloc(c, A.pl_cnt, n) | dalloc(c, A.disp_p, 3 * (size_t)n);
  r |= dalloc(c, c->handled_s, n) | dalloc(c, c->handled_l, n);
  {  // reduction tiles
    std::vector<int2> rb;
    std::vector<int> ro(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
      ro[s] = (int)rb.size();
      int m = off[s + 1] - off[s];
      for (int t = 0; t < std::max(1, (m + kRedTile - 1) / kRedTile); t++)
        rb.push_back(make_int2(s, off[s] + t * kRedTile));
    }
    ro[nsys] = (int)rb.size();
    c->nred = (int)rb.size();
    r |= dalloc(c, c->rb, rb.size()) | dalloc(c, c->rb_off, nsys + 1) | dalloc(c, c->red_part, 8 * rb.size()) |
         dalloc(c, c->simt, nsys) | dalloc(c, c->tfin, nsys) | dalloc(c, c->dtprev, nsys) |
         dalloc(c, c->iter, nsys);
    c->traj_on = false;
    c->tb = TrajBuf{};
    c->frames_h.clear();
    c->recs_h.clear();
    if (!r) {
      HIPCHK(c, hipMemcpy(c->rb, rb.data(), rb.size() * sizeof(int2), hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->rb_off, ro.data(), (nsys + 1) * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  A.ml = nullptr;
  A.ml_cnt = nullptr;
  A.disp_e = nullptr;
  c->eflags = nullptr;
  if (c->slot_eam >= 0 || c->slot_tb >= 0)
    r |= dalloc(c, A.ml, (size_t)n * kMetalRow) | dalloc(c, A.ml_cnt, n) | dalloc(c, A.disp_e, 3 * (size_t)n) |
         dalloc(c, c->eflags, nsys);
  for (int k = 0; k < 2; k++) {
    auto& E = c->emb[k];
    E = mdg_ctx::EmbedSet{};
    if ((k == 0 ? c->slot_eam : c->slot_tb) >= 0)
      r |= dalloc(c, E.e, n) | dalloc(c, E.fe, n) | dalloc(c, E.row, (size_t)n * kEamRow) | dalloc(c, E.cnt, n);
  }
  if (r) return 1;
  r |= dalloc(c, c->cell_of_s, n) | dalloc(c, c->atoms_s, n) | dalloc(c, c->cell_of_l, n) |
       dalloc(c, c->atoms_l, n);
  if (r || setup_grids(c, p)) return 1;
  std::vector<double> xs(n), ys(n), zs2(n), vxs(n, 0.0), vys(n, 0.0), vzs(n, 0.0);
  std::vector<int> px(n), py(n), pz(n);
  for (int i = 0; i < n; i++) {
    xs[i] = p[3 * i];
    ys[i] = p[3 * i + 1];
    zs2[i] = p[3 * i + 2];
    if (vel && !fx[i]) {
      vxs[i] = vel[3 * i];
      vys[i] = vel[3 * i + 1];
      vzs[i] = vel[3 * i + 2];
    }
    px[i] = pc[3 * i];
    py[i] = pc[3 * i + 1];
    pz[i] = pc[3 * i + 2];
  }
  auto up = [&](void* d, const void* s, size_t b) { return hipMemcpy(d, s, b, hipMemcpyHostToDevice); };
  HIPCHK(c, up(Ld, c->Lh.data(), 3 * (size_t)nsys * 8));
  HIPCHK(c, up(sysd, sy.data(), (size_t)n * 4));
  HIPCHK(c, up(A.sys_off, off.data(), (nsys + 1) * 4));
  HIPCHK(c, up(A.x, xs.data(), (size_t)n * 8));
  HIPCHK(c, up(A.y, ys.data(), (size_t)n * 8));
  HIPCHK(c, up(A.z, zs2.data(), (size_t)n * 8));
  HIPCHK(c, up(A.vx, vxs.data(), (size_t)n * 8));
  HIPCHK(c, up(A.vy, vys.data(), (size_t)n * 8));
  HIPCHK(c, up(A.vz, vzs.data(), (size_t)n * 8));
  HIPCHK(c, hipMemset(A.ax, 0, (size_t)n * 8));
  HIPCHK(c, hipMemset(A.ay, 0, (size_t)n * 8));
  HIPCHK(c, hipMemset(A.az, 0, (size_t)n * 8));
  HIPCHK(c, up(A.mass, ms.data(), (size_t)n * 8));
  HIPCHK(c, up(A.Z, zs.data(), (size_t)n * 8));
  HIPCHK(c, up(A.type, ty.data(), (size_t)n * 4));
  HIPCHK(c, up(A.elem, el.data(), (size_t)n * 4));
  HIPCHK(c, up(A.fixed, fx.data(), n));
  HIPCHK(c, up(A.pcx, px.data(), (size_t)n * 4));
  HIPCHK(c, up(A.pcy, py.data(), (size_t)n * 4));
  HIPCHK(c, up(A.pcz, pz.data(), (size_t)n * 4));
  k_pack_pos<<<blocks(n), 256, 0, c->stream>>>(A);  // after the positions and types are uploaded
  HIPCHK(c, hipMemset(A.disp_s, 0, 3 * (size_t)n * 8));
  HIPCHK(c, hipMemset(A.disp_l, 0, 3 * (size_t)n * 8));
  if (A.disp_p) HIPCHK(c, hipMemset(A.disp_p, 0, 3 * (size_t)n * 8));
  if (A.disp_e) HIPCHK(c, hipMemset(A.disp_e, 0, 3 * (size_t)n * 8));
  if (A.ml_cnt) HIPCHK(c, hipMemset(A.ml_cnt, 0xff, (size_t)n * 4));  // -1: no metal row yet
  if (A.pl_cnt) HIPCHK(c, hipMemset(A.pl_cnt, 0xff, (size_t)n * 4));  // -1: no pair row yet
  HIPCHK(c, hipMemset(A.nls_cnt, 0, (size_t)n * 4));
  HIPCHK(c, hipMemset(A.nll_cnt, 0, (size_t)n * 4));
  if (A.nll_split) HIPCHK(c, hipMemset(A.nll_split, 0, (size_t)n * 4));
  if (A.nbl) HIPCHK(c, hipMemset(A.nbl, 0, (size_t)n));
  HIPCHK(c, hipMemset(A.bcnt, 0, (size_t)n * 4));
  HIPCHK(c, up(c->handled_s, hs.data(), n));
  HIPCHK(c, up(c->handled_l, hl.data(), n));
  for (int k = 0; k < 2; k++)
    if (c->emb[k].e) {  // Ackland: Cu 0, Ag 1, Au 2; TightBinding: Cu 0
      std::vector<signed char> ek(n);
      for (int i = 0; i < n; i++)
        ek[i] = el[i] == 64 ? 0 : (k == 0 && el[i] == 108) ? 1 : (k == 0 && el[i] == 197) ? 2 : -1;
      HIPCHK(c, up(c->emb[k].e, ek.data(), n));
    }
  int fl[4] = {1, 1, 1, 1};
  std::vector<int> sfl(2 * (size_t)nsys, 1);
  HIPCHK(c, up(c->flags, fl, sizeof(fl)));
  if (c->eflags) HIPCHK(c, up(c->eflags, sfl.data(), (size_t)nsys * 4));
  HIPCHK(c, up(c->sflags, sfl.data(), sfl.size() * 4));
  HIPCHK(c, up(c->pflags, sfl.data(), (size_t)nsys * 4));
  HIPCHK(c, up(c->dtp, c->dt_host.data(), (size_t)nsys * 8));
  {  // simTime 0, no final time (SimLoop.cxx:46)
    std::vector<double> t0(nsys, 0.0), tf(nsys, 1e300);
    std::vector<long long> it0(nsys, 0);
    HIPCHK(c, up(c->simt, t0.data(), (size_t)nsys * 8));
    HIPCHK(c, up(c->tfin, tf.data(), (size_t)nsys * 8));
    HIPCHK(c, up(c->dtprev, c->dt_host.data(), (size_t)nsys * 8));  // dt_prev = dt (SimLoop.cxx:248)
    HIPCHK(c, up(c->iter, it0.data(), (size_t)nsys * 8));
  }
  c->iteration = 0;
  c->acc_keep = nullptr;
  return 0;
}

int mdg_set_system(mdg_ctx* c, int n, const int32_t* elem, const double* xyz, const double* vel,
                   const double pbc[3], const uint8_t* fixed) {
  if (n <= 0) {
    c->err = "empty system (AtomsContainer.cxx:108)";
    return 1;
  }
  int32_t cnt = n;
  return mdg_set_systems(c, 1, &cnt, elem, xyz, vel, pbc, fixed);
}

int mdg_num_systems(const mdg_ctx* c) { return c->nsys; }


// The PBC the thermal bath's geometry REQUIREs of every atom it applies to
// (SimLoop.hpp:141 lateral PBC for the box, :166 none for the sphere), each
// atom with its own PBC vector.
int bath_pbc_check(mdg_ctx* c, int s, int geom) {
  if (geom != MDG_TB_BOX && geom != MDG_TB_SPHERE) return 0;
  const bool mixed = !c->apbc_h.empty() && c->apbc_h[s];
  for (int i = c->off_h[s]; i < (mixed ? c->off_h[s + 1] : c->off_h[s] + 1); i++) {
    if (mixed && !c->tb_apply_h.empty() && !c->tb_apply_h[i]) continue;
    double L[3];
    for (int d = 0; d < 3; d++)
      L[d] = (mixed && ((c->open_h[i] >> d) & 1)) ? kNoPbc : c->Lh[3 * s + d];
    if (geom == MDG_TB_BOX && (L[0] == kNoPbc || L[1] == kNoPbc)) {
      c->err = "thermal bath box needs lateral PBC (SimLoop.hpp:141) [system " + std::to_string(s) + "]";
      return 1;
    }
    if (geom == MDG_TB_SPHERE && (L[0] != kNoPbc || L[1] != kNoPbc || L[2] != kNoPbc)) {
      c->err = "thermal bath sphere needs open boundaries (SimLoop.hpp:166) [system " + std::to_string(s) + "]";
      return 1;
    }
  }
  return 0;
}

int mdg_set_thermal_bath(mdg_ctx* c, int system, const mdg_thermal_bath* tb) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (!tb || system < -1 || system >= c->nsys) {
    c->err = "mdg_set_thermal_bath: bad system index or null parameters";
    return 1;
  }
  if (tb->geometry < MDG_TB_NONE || tb->geometry > MDG_TB_SPHERE || !std::isfinite(tb->To) ||
      !std::isfinite(tb->gamma)) {
    c->err = "mdg_set_thermal_bath: bad geometry or non-finite To/gamma";
    return 1;
  }
  int s0 = system < 0 ? 0 : system, s1 = system < 0 ? c->nsys : system + 1;
  for (int s = s0; s < s1; s++)
    if (bath_pbc_check(c, s, tb->geometry)) return 1;
  if (bath_alloc(c)) return 1;
  for (int s = s0; s < s1; s++)
    c->bath_h[s] = Bath{tb->geometry, tb->To,        tb->gamma,      tb->z_min,  tb->d_boundary, tb->z_min_free,
                        tb->center[0], tb->center[1], tb->center[2], tb->radius};
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->bath_d, c->bath_h.data(), c->nsys * sizeof(Bath), hipMemcpyHostToDevice));
  c->bath_on = false;
  for (const Bath& b : c->bath_h) c->bath_on = c->bath_on || b.geom != TB_NONE;
  drop_graphs(c);
  return 0;
}

int mdg_set_bath_atoms(mdg_ctx* c, const uint8_t* apply) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (bath_alloc(c)) return 1;
  std::vector<unsigned char> a(c->n, 1);
  if (apply)
    for (int i = 0; i < c->n; i++) a[i] = apply[i] ? 1 : 0;
  std::vector<unsigned char> prev = c->tb_apply_h;
  c->tb_apply_h = a;
  for (int s = 0; s < c->nsys; s++)
    if (bath_pbc_check(c, s, c->bath_h[s].geom)) {
      c->tb_apply_h = prev;
      return 1;
    }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->tb_apply_d, a.data(), c->n, hipMemcpyHostToDevice));
  return 0;
}

int mdg_set_time(mdg_ctx* c, int system, double sim_time, double sim_time_final, double dt) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (system < -1 || system >= c->nsys) {
    c->err = "bad system index";
    return 1;
  }
  if (!(dt > 0.0)) {
    c->err = "dt must be positive";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int s0 = system < 0 ? 0 : system, s1 = system < 0 ? c->nsys : system + 1;
  std::vector<double> t(s1 - s0, sim_time), tf(s1 - s0, sim_time_final), d(s1 - s0, dt);
  HIPCHK(c, hipMemcpy(c->simt + s0, t.data(), t.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->tfin + s0, tf.data(), tf.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->dtp + s0, d.data(), d.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->dtprev + s0, d.data(), d.size() * 8, hipMemcpyHostToDevice));  // SimLoop.cxx:248
  for (int s = s0; s < s1; s++) c->dt_host[s] = dt;
  return 0;
}

int mdg_get_time(mdg_ctx* c, double* out) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> t(c->nsys), tf(c->nsys), d(c->nsys);
  HIPCHK(c, hipMemcpy(t.data(), c->simt, c->nsys * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(tf.data(), c->tfin, c->nsys * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(d.data(), c->dtp, c->nsys * 8, hipMemcpyDeviceToHost));
  for (int s = 0; s < c->nsys; s++) {
    out[3 * s] = t[s];
    out[3 * s + 1] = tf[s];
    out[3 * s + 2] = d[s];
  }
  return 0;
}

int mdg_set_trajectory_output(mdg_ctx* c, const uint8_t* snapshot_atoms, int xva_frames) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);  // the step's graph gains or loses the capture kernel
  c->traj_on = false;
  dfree(c, c->tb.rec);
  dfree(c, c->tb.frm);
  dfree(c, c->tb.cnt);
  dfree(c, c->tb.sel);
  dfree(c, c->tb.sel_off);
  c->tb = TrajBuf{};
  c->frames_h.clear();
  c->recs_h.clear();
  if (!snapshot_atoms && !xva_frames) return 0;
  const int nsys = c->nsys, n = c->n;
  std::vector<int> sel, off(nsys + 1, 0);
  for (int s = 0; s < nsys; s++) {
    off[s] = (int)sel.size();
    if (snapshot_atoms)
      for (int i = c->off_h[s]; i < c->off_h[s + 1]; i++)
        if (snapshot_atoms[i]) sel.push_back(i);
  }
  off[nsys] = (int)sel.size();
  TrajBuf& t = c->tb;
  t.cap_frm = 2LL * kTrajChunk * nsys + 64;
  t.cap_rec = (long long)kTrajChunk * (long long)sel.size() + (xva_frames ? 2LL * n : 0) + 64;
  int* seld = nullptr;
  int* offd = nullptr;
  if (dalloc(c, t.rec, (size_t)t.cap_rec) || dalloc(c, t.frm, (size_t)t.cap_frm) || dalloc(c, t.cnt, 3))
    return 1;
  if (!sel.empty()) {
    if (dalloc(c, seld, sel.size()) || dalloc(c, offd, nsys + 1)) return 1;
    HIPCHK(c, hipMemcpy(seld, sel.data(), sel.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(offd, off.data(), (nsys + 1) * 4, hipMemcpyHostToDevice));
  }
  t.sel = seld;
  t.sel_off = offd;
  t.xva = xva_frames ? 1 : 0;
  c->traj_on = true;
  return 0;
}

int mdg_trajectory_size(mdg_ctx* c, int64_t* nframes, int64_t* nrecords) {
  if (nframes) *nframes = (int64_t)c->frames_h.size();
  if (nrecords) *nrecords = (int64_t)c->recs_h.size();
  return 0;
}

int mdg_take_trajectory(mdg_ctx* c, mdg_frame* frames, mdg_atom_record* records) {
  if (frames && !c->frames_h.empty()) std::memcpy(frames, c->frames_h.data(), c->frames_h.size() * sizeof(mdg_frame));
  if (records && !c->recs_h.empty())
    std::memcpy(records, c->recs_h.data(), c->recs_h.size() * sizeof(mdg_atom_record));
  c->frames_h.clear();
  c->recs_h.clear();
  return 0;
}

int mdg_bath_energy(mdg_ctx* c, double* out) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!c->ebath) {
    for (int s = 0; s < c->nsys; s++) out[s] = 0.0;
    return 0;
  }
  HIPCHK(c, hipMemcpy(out, c->ebath, c->nsys * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int mdg_energy_grad(mdg_ctx* c, double* e_terms, double* grad) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  force_call(c);
  int nt = (int)c->kinds.size();
  k_sys_energy<<<c->nsys, 256, 0, c->stream>>>(nt, c->A.sys_off, c->n, c->eatom, c->eacc);
  if (check_err(c)) return 1;
  std::vector<double> e(8 * (size_t)c->nsys);
  HIPCHK(c, hipMemcpyAsync(e.data(), c->eacc, e.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (grad)
    HIPCHK(c, hipMemcpyAsync(grad, c->A.grad, 3 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (e_terms)
    for (int s = 0; s < c->nsys; s++)
      for (int t = 0; t < nt; t++) e_terms[(size_t)s * nt + t] = e[8 * (size_t)s + t];
  return 0;
}

}  // extern "C"

namespace {
// Static scratch of mdg_energy_grad_noalloc: every device buffer of that call
// (atoms, cell grids, lists, bond records, the b* queue, tables, counters)
// is carved from this array; each GPU holds its own copy, at its own
// address, until the process ends.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr int kNoAllocMaxAtoms = 4096;
constexpr int kMaxDevices = 64;
__device__ __align__(256) char scratch[kNoAllocBytes];
std::mutex g_scratch_mu;
char* g_scratch_at[kMaxDevices] = {};

// The device address of the current device's scratch, looked up once per
// device (the array's name is only a host-side handle).
int scratch_of(int dev, char*& out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_at[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return 1;
    }
    g_scratch_at[dev] = static_cast<char*>(p);
  }
  out = g_scratch_at[dev];
  return 0;
}
}  // namespace

extern "C" {

int mdg_energy_grad_noalloc(const mdg_config* cfg, int n, const int32_t* elem, const double* xyz,
                            const double pbc[3], double* e_terms, double* grad) {
  if (n <= 0 || n > kNoAllocMaxAtoms) {
    g_create_err = "mdg_energy_grad_noalloc: n = " + std::to_string(n) + " outside 1.." +
                   std::to_string(kNoAllocMaxAtoms) + " (the static scratch's limit)";
    return 1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    g_create_err = "mdg_energy_grad_noalloc: no current HIP device";
    return 1;
  }
  char* base = nullptr;
  if (scratch_of(dev, base)) return 1;
  mdg_ctx c;
  c.arena = base;
  c.arena_cap = kNoAllocBytes;
  int rc = ctx_init(&c, cfg);
  if (!rc) rc = mdg_set_system(&c, n, elem, xyz, nullptr, pbc, nullptr);
  if (!rc) rc = mdg_energy_grad(&c, e_terms, grad);
  if (rc) g_create_err = "mdg_energy_grad_noalloc: " + c.err;
  return rc;
}

int mdg_system_errors(mdg_ctx* c, uint32_t* out) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (check_err(c, true)) return 1;  // errors of global resources fail here
  for (int s = 0; s < c->nsys; s++) out[s] = c->failed[s];
  std::fill(c->failed.begin(), c->failed.end(), 0u);
  HIPCHK(c, hipMemset(c->derr, 0, (1 + (size_t)c->nsys) * sizeof(unsigned)));
  return 0;
}

int mdg_set_virial(mdg_ctx* c, int on) {
  if (c->n == 0) {
    c->err = "mdg_set_virial: set the system(s) first";
    return 1;
  }
  if (on && !c->vir && (dalloc(c, c->vir, 6 * (size_t)c->nsys) || dalloc(c, c->virfix, 12 * (size_t)c->nsys)))
    return 1;
  if ((on != 0) != c->virial) drop_graphs(c);  // the captured step's launches change
  c->virial = on != 0;
  return 0;
}

int mdg_virial(mdg_ctx* c, double* out) {
  if (!c->virial) {
    c->err = "mdg_virial: the virial is off (mdg_set_virial)";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(out, c->vir, 6 * (size_t)c->nsys * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int mdg_step_virial(mdg_ctx* c, double* out) {
  if (!c->virial) {
    c->err = "mdg_step_virial: the virial is off (mdg_set_virial)";
    return 1;
  }
  std::copy(c->vir_steps.begin(), c->vir_steps.end(), out);
  return 0;
}

int mdg_get_params(const mdg_ctx* c, double* theta) {
  Tables t = c->host_T;
  theta_io(t, theta, false);
  return 0;
}

int mdg_set_params(mdg_ctx* c, const double* theta) {
  std::vector<double> th(theta, theta + kNumTheta);
  for (double v : th)
    if (!std::isfinite(v)) {
      c->err = "mdg_set_params: non-finite parameter";
      return 1;
    }
  Tables t = c->host_T;
  theta_io(t, th.data(), true);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      if (!(t.lj.R[a][b][0] > t.lj.RLJ[a][b][1]) || !(t.lj.sigma[a][b] > 0)) {
        c->err = "mdg_set_params: 2^(1/6) sigma must stay below the LJ switch start (AIREBO_LJ.cxx:254-257)";
        return 1;
      }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->dT, &t, sizeof(Tables), hipMemcpyHostToDevice));
  if (c->mixed) {
    TablesT<float> hf;
    convert_tables(t, hf);
    HIPCHK(c, hipMemcpy(c->dTf, &hf, sizeof(hf), hipMemcpyHostToDevice));
  }
  c->host_T = t;
  return mark_pair_rows(c);  // the LJ cutoffs bound the pair rows
}

int mdg_energy_grad_params(mdg_ctx* c, double* e_terms, double* grad, double* dtheta) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (c->mixed) {
    c->err = "parameter gradients are computed by the FP64 form (create the context with MDG_PREC_FP64)";
    return 1;
  }
  if (!c->dtheta && dalloc(c, c->dtheta, (size_t)kNumTheta * c->nsys)) return 1;
  k_zero<<<blocks((long long)kNumTheta * c->nsys), 256, 0, c->stream>>>(c->dtheta, (size_t)kNumTheta * c->nsys);
  c->theta_mode = true;
  int rc = mdg_energy_grad(c, e_terms, grad);
  c->theta_mode = false;
  if (rc) return rc;
  if (dtheta)
    HIPCHK(c, hipMemcpy(dtheta, c->dtheta, (size_t)kNumTheta * c->nsys * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int mdg_nl_export(mdg_ctx* c, int term, int32_t* counts, int32_t* idx, int64_t cap) {
  if (term < 0 || term >= (int)c->kinds.size()) {
    c->err = "bad term";
    return 1;
  }
  int kind = c->kinds[term];
  bool shortlist = (kind == MDG_POT_REBO || kind == MDG_POT_BRENNER || kind == MDG_POT_ETORS);
  int n = c->n;
  // make sure lists reflect the current positions' rebuild state
  nl_build(c, shortlist);
  k_clear_list<<<blocks(c->nsys), 256, 0, c->stream>>>(c->nsys, shortlist ? 0 : 1, c->flags, c->sflags);
  if (check_err(c)) return 1;
  if (!shortlist && mark_pair_rows(c)) return 1;  // rows are filtered from the long list just rebuilt
  int stride = shortlist ? c->A.cap_s : c->A.cap_l;
  std::vector<int> cnt(n), lst((size_t)n * stride), ty(n), el(n);
  HIPCHK(c, hipMemcpy(cnt.data(), shortlist ? c->A.nls_cnt : c->A.nll_cnt, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(lst.data(), shortlist ? c->A.nls : c->A.nll, (size_t)n * stride * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ty.data(), c->A.type, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(el.data(), c->A.elem, n * 4, hipMemcpyDeviceToHost));
  auto zbl_handled = [](int e) { return e == 1 || e == 12 || e == 64 || e == 197 || e == 108 || e == 40 || e == 131; };
  auto flj_handled = [](int e) { return e == 1 || e == 12 || e == 64 || e == 197; };  // FLJ.cxx:38-52
  auto eam_handled = [](int e) { return e == 64 || e == 108 || e == 197; };             // Ackland.cxx:236-262
  auto fbm_handled = [](int e) { return e == 64 || e == 40; };                          // FBM.cxx:37-38
  auto tb_handled = [](int e) { return e == 64; };                                      // TightBinding.cxx:163
  int64_t pos = 0;
  for (int i = 0; i < n; i++) {
    int m = 0;
    bool row_ok = true;
    if (kind == MDG_POT_AIREBO) row_ok = ty[i] <= T_H;
    if (kind == MDG_POT_FBZL) row_ok = zbl_handled(el[i]);
    if (kind == MDG_POT_FLJ) row_ok = flj_handled(el[i]);
    if (kind == MDG_POT_ACKLAND) row_ok = eam_handled(el[i]);
    if (kind == MDG_POT_FBM) row_ok = fbm_handled(el[i]);
    if (kind == MDG_POT_TB) row_ok = tb_handled(el[i]);
    if (row_ok)
      for (int q = 0; q < cnt[i]; q++) {
        int j = lst[(size_t)i * stride + q];
        if (kind == MDG_POT_AIREBO && ty[j] > T_H) continue;
        if (kind == MDG_POT_FBZL && !zbl_handled(el[j])) continue;
        if (kind == MDG_POT_FLJ && !flj_handled(el[j])) continue;
        if (kind == MDG_POT_ACKLAND && !eam_handled(el[j])) continue;
        if (kind == MDG_POT_FBM && !fbm_handled(el[j])) continue;
        if (kind == MDG_POT_TB && !tb_handled(el[j])) continue;
        if (pos >= cap) {
          c->err = "nl_export capacity";
          return 1;
        }
        idx[pos++] = j;
        m++;
      }
    // the mixed form's long rows hold the j < i, then the j > i entries in
    // candidate order; the reference's lists are ascending
    if (!shortlist) std::sort(idx + (pos - m), idx + pos);
    counts[i] = m;
  }
  return 0;
}

int mdg_step_nve(mdg_ctx* c, int nsteps, double fixed_dt, double* epot, double* ekin, double* dt_used,
                 double* netforce) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (nsteps <= 0) return 0;
  Atoms& A = c->A;
  int n = c->n, nsys = c->nsys;
  const std::vector<unsigned> failed0 = c->failed;
  // per-step, per-system outputs are staged in chunks of at most 2^20 entries
  // with trajectory output, at most kTrajChunk steps between drains of its buffers
  const int chunk = std::max(1, std::min(c->traj_on ? kTrajChunk : 4096, (1 << 20) / nsys));
  if (c->step_cap < chunk) {
    if (dalloc(c, c->stepout, (size_t)chunk * nsys * 4)) return 1;
    if (!c->stepidx && dalloc(c, c->stepidx, 1)) return 1;
    c->step_cap = chunk;
  }
  if (c->virial && c->stepvir_cap < chunk) {
    if (dalloc(c, c->stepvir, (size_t)chunk * nsys * 6)) return 1;
    c->stepvir_cap = chunk;
  }
  c->vir_steps.clear();
  int fixed = fixed_dt > 0 ? 1 : 0;
  if (fixed) k_fill<<<blocks(nsys), 256, 0, c->stream>>>(c->dtp, nsys, fixed_dt);
  int nterms = (int)c->kinds.size();
  if (c->iteration == 0) {  // SimLoop.cxx:305-321, for the systems at their iteration 0
    force_call(c);
    launch(c, "k_init_acc", [&] { k_init_acc<<<blocks(n), 256, 0, c->stream>>>(A, c->acc_keep); });
  }
  double lim_s = 0.44 * c->rc_short, lim_l = 0.44 * c->rc_long;
  if (!c->have_short) lim_s = 1e300;
  if (!c->have_long) lim_l = 1e300;
  // One iteration of the loop body for every system of the batch; every size
  // and flag it needs lives in device memory, so the same launches can be
  // captured once and replayed.
  auto body = [&]() {
    if (c->traj_on)
      launch(c, "k_traj_capture", [&] {
        k_traj_capture<<<nsys, 256, 0, c->stream>>>(A, c->tb, c->simt, c->tfin, c->dtprev, c->iter, c->dtp);
      });
    if (c->bath_on)
      launch(c, "k_bath_temp", [&] {
        k_bath_temp_part<<<c->nred, 256, 0, c->stream>>>(A, c->box, c->rb, c->red_part);
        k_bath_temp_merge<<<nsys, 64, 0, c->stream>>>(A, c->rb_off, c->red_part);
      });
    launch(c, "k_drift", [&] {
      k_drift_tile<<<c->nred, 256, 0, c->stream>>>(A, c->rb, c->dtp, c->simt, c->tfin, c->red_part);
    });
    launch(c, "k_sys_disp", [&] {
      k_sys_disp_merge<<<nsys, 64, 0, c->stream>>>(c->rb_off, c->red_part, lim_s, lim_l, c->pair_margin,
                                                   A.disp_p ? 1 : 0, c->metal_skin, A.disp_e ? 1 : 0, c->sflags,
                                                   c->pflags, c->eflags, c->flags, c->counters);
    });
    force_call(c);
    launch(c, "k_kick", [&] {
      k_kick_tile<<<c->nred, 256, 0, c->stream>>>(A, c->box, c->rb, c->dtp, c->simt, c->tfin, nterms, c->eatom,
                                                  c->red_part);
    });
    launch(c, "k_sys_step_end", [&] {
      k_sys_step_merge<<<nsys, 64, 0, c->stream>>>(c->rb_off, c->red_part, c->dtp, fixed, c->stepout, c->stepidx,
                                                   nsys, A.tb_dE ? c->ebath : nullptr, c->simt, c->tfin, c->dtprev,
                                                   c->iter, c->derr + 1);
    });
    if (c->virial) k_step_virial<<<blocks(6LL * nsys), 256, 0, c->stream>>>(nsys, c->vir, c->stepvir, c->stepidx);
    k_step_advance<<<1, 64, 0, c->stream>>>(c->stepidx);
  };
  bool use_graph = !c->profiling && c->use_graphs;
  // The whole step lives on the device, so `reps` steps capture into one
  // graph: one launch advances every system of the batch by kGraphBlock
  // steps, the rest of a chunk goes one step per launch.
  auto capture = [&](int idx, int reps) -> int {
    if (c->graph_exec[idx]) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < reps; r++) body();
    HIPCHK(c, hipStreamEndCapture(c->stream, &g));
    HIPCHK(c, hipGraphInstantiate(&c->graph_exec[idx], g, nullptr, nullptr, 0));
    c->graphs.push_back(g);
    return 0;
  };
  const int block = c->graph_block;
  std::vector<double> o((size_t)chunk * nsys * 4);
  for (int s0 = 0; s0 < nsteps; s0 += chunk) {
    int m = std::min(chunk, nsteps - s0);
    k_zero_i<<<1, 64, 0, c->stream>>>(c->stepidx, 1);
    if (c->traj_on) k_zero<<<1, 64, 0, c->stream>>>(reinterpret_cast<double*>(c->tb.cnt), 3);  // all-zero bits
    int s = 0;
    if (use_graph) {
      if (block > 1 && m >= block && capture(fixed + 2, block)) return 1;
      for (; block > 1 && s + block <= m; s += block) HIPCHK(c, hipGraphLaunch(c->graph_exec[fixed + 2], c->stream));
      if (s < m && capture(fixed, 1)) return 1;
      for (; s < m; s++) HIPCHK(c, hipGraphLaunch(c->graph_exec[fixed], c->stream));
    } else {
      for (; s < m; s++) body();
    }
    if (check_err(c, true)) return 1;  // systems with their own errors stand still, the rest go on
    if (c->traj_on && drain_trajectory(c)) return 1;
    HIPCHK(c, hipMemcpy(o.data(), c->stepout, (size_t)m * nsys * 4 * sizeof(double), hipMemcpyDeviceToHost));
    if (c->virial) {
      size_t v0 = c->vir_steps.size();
      c->vir_steps.resize(v0 + (size_t)m * nsys * 6);
      HIPCHK(c, hipMemcpy(&c->vir_steps[v0], c->stepvir, (size_t)m * nsys * 6 * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    for (int s = 0; s < m; s++)
      for (int q = 0; q < nsys; q++) {
        const double* v = &o[((size_t)s * nsys + q) * 4];
        size_t at = (size_t)(s0 + s) * nsys + q;
        if (epot) epot[at] = v[0];
        if (ekin) ekin[at] = v[1];
        if (dt_used) dt_used[at] = v[2];
        if (netforce) netforce[at] = v[3];
      }
  }
  c->iteration += nsteps;
  for (int s = 0; s < nsys; s++)
    if (c->failed[s] & ~failed0[s]) {  // c->err names the systems and errors
      c->err = "some systems stopped on device errors (the others ran every step):" + c->err.substr(13);
      return MDG_SYSTEMS_STOPPED;
    }
  return 0;
}

int mdg_get_state(mdg_ctx* c, double* xyz, double* vel, double* acc, int32_t* pbc_count) {
  int n = c->n;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> a(n), b(n), d(n);
  auto get3 = [&](double* out, double* X, double* Y, double* Z) -> int {
    if (!out) return 0;
    HIPCHK(c, hipMemcpy(a.data(), X, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(b.data(), Y, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(d.data(), Z, n * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
      out[3 * i] = a[i];
      out[3 * i + 1] = b[i];
      out[3 * i + 2] = d[i];
    }
    return 0;
  };
  if (get3(xyz, c->A.x, c->A.y, c->A.z) || get3(vel, c->A.vx, c->A.vy, c->A.vz) || get3(acc, c->A.ax, c->A.ay, c->A.az))
    return 1;
  if (pbc_count) {
    std::vector<int> q(n);
    int* src[3] = {c->A.pcx, c->A.pcy, c->A.pcz};
    for (int dd = 0; dd < 3; dd++) {
      HIPCHK(c, hipMemcpy(q.data(), src[dd], n * 4, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; i++) pbc_count[3 * i + dd] = q[i];
    }
  }
  return 0;
}

int mdg_set_accelerations(mdg_ctx* c, const double* acc, const uint8_t* keep) {
  int n = c->n;
  if (n == 0) {
    c->err = "no system";
    return 1;
  }
  if (c->iteration != 0) {
    c->err = "mdg_set_accelerations: only before the first mdg_step_nve of the systems";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> a(n), b(n), d(n), nt(3 * (size_t)n);
  std::vector<unsigned char> fx(n);
  HIPCHK(c, hipMemcpy(fx.data(), c->A.fixed, n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    const bool f = fx[i] != 0;  // a fixed atom keeps a = 0 (Atom::fix, Atom.hpp:58)
    a[i] = f ? 0.0 : acc[3 * i];
    b[i] = f ? 0.0 : acc[3 * i + 1];
    d[i] = f ? 0.0 : acc[3 * i + 2];
    nt[3 * i] = a[i];
    nt[3 * i + 1] = b[i];
    nt[3 * i + 2] = d[i];
  }
  HIPCHK(c, hipMemcpy(c->A.ax, a.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->A.ay, b.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->A.az, d.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->A.an_nt, nt.data(), nt.size() * 8, hipMemcpyHostToDevice));
  std::vector<unsigned char> k(c->nsys, 1);
  if (keep)
    for (int s = 0; s < c->nsys; s++) k[s] = keep[s] ? 1 : 0;
  if (!c->acc_keep && dalloc(c, c->acc_keep, c->nsys)) return 1;
  HIPCHK(c, hipMemcpy(c->acc_keep, k.data(), c->nsys, hipMemcpyHostToDevice));
  return 0;
}

int mdg_set_iteration(mdg_ctx* c, int system, int64_t iteration) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (system < -1 || system >= c->nsys || iteration < 0) {
    c->err = "mdg_set_iteration: bad system index or negative iteration";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int s0 = system < 0 ? 0 : system, s1 = system < 0 ? c->nsys : system + 1;
  std::vector<long long> it(s1 - s0, (long long)iteration);
  HIPCHK(c, hipMemcpy(c->iter + s0, it.data(), it.size() * 8, hipMemcpyHostToDevice));
  return 0;
}

int mdg_get_iteration(mdg_ctx* c, int64_t* out) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<long long> it(c->nsys);
  HIPCHK(c, hipMemcpy(it.data(), c->iter, it.size() * 8, hipMemcpyDeviceToHost));
  for (int s = 0; s < c->nsys; s++) out[s] = it[s];
  return 0;
}

int mdg_set_positions(mdg_ctx* c, const double* xyz, const int32_t* pbc_count, const double* dr) {
  int n = c->n;
  if (n == 0) {
    c->err = "no system";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> a(n), b(n), d(n);
  for (int i = 0; i < n; i++) {
    a[i] = xyz[3 * i];
    b[i] = xyz[3 * i + 1];
    d[i] = xyz[3 * i + 2];
  }
  HIPCHK(c, hipMemcpy(c->A.x, a.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->A.y, b.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->A.z, d.data(), n * 8, hipMemcpyHostToDevice));
  k_pack_pos<<<blocks(n), 256, 0, c->stream>>>(c->A);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (pbc_count) {
    std::vector<int> q(n);
    int* dst[3] = {c->A.pcx, c->A.pcy, c->A.pcz};
    for (int dd = 0; dd < 3; dd++) {
      for (int i = 0; i < n; i++) q[i] = pbc_count[3 * i + dd];
      HIPCHK(c, hipMemcpy(dst[dd], q.data(), n * 4, hipMemcpyHostToDevice));
    }
  }
  if (!dr) {
    int fl[4] = {1, 1, 1, 1};
    std::vector<int> sfl(2 * (size_t)c->nsys, 1);
    HIPCHK(c, hipMemcpy(c->flags, fl, sizeof(fl), hipMemcpyHostToDevice));
    if (c->eflags) HIPCHK(c, hipMemcpy(c->eflags, sfl.data(), (size_t)c->nsys * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->sflags, sfl.data(), sfl.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->pflags, sfl.data(), (size_t)c->nsys * 4, hipMemcpyHostToDevice));
    return 0;
  }
  // accumulate dr into both lists' displacements, then the skin test
  std::vector<double> s((size_t)3 * n), l((size_t)3 * n);
  HIPCHK(c, hipMemcpy(s.data(), c->A.disp_s, 3 * (size_t)n * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(l.data(), c->A.disp_l, 3 * (size_t)n * 8, hipMemcpyDeviceToHost));
  for (size_t q = 0; q < (size_t)3 * n; q++) {
    s[q] += dr[q];
    l[q] += dr[q];
  }
  HIPCHK(c, hipMemcpy(c->A.disp_s, s.data(), 3 * (size_t)n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->A.disp_l, l.data(), 3 * (size_t)n * 8, hipMemcpyHostToDevice));
  for (double* dd : {c->A.disp_p, c->A.disp_e})
    if (dd) {
      HIPCHK(c, hipMemcpy(s.data(), dd, 3 * (size_t)n * 8, hipMemcpyDeviceToHost));
      for (size_t q = 0; q < (size_t)3 * n; q++) s[q] += dr[q];
      HIPCHK(c, hipMemcpy(dd, s.data(), 3 * (size_t)n * 8, hipMemcpyHostToDevice));
    }
  double lim_s = c->have_short ? 0.44 * c->rc_short : 1e300;
  double lim_l = c->have_long ? 0.44 * c->rc_long : 1e300;
  sys_disp(c, lim_s, lim_l);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mdg_set_atom_pbc(mdg_ctx* c, const double* pbc, const double* xyz, const int32_t* pbc_count) {
  const int n = c->n, nsys = c->nsys;
  if (n == 0) {
    c->err = "no system";
    return 1;
  }
  if (!pbc || !xyz) {
    c->err = "mdg_set_atom_pbc: null pbc or positions";
    return 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // every list's minimum image must be the reference's depos for the pairs it
  // lists: L >= 2 (1 + 0.44) Rc on the axes where a system's atoms differ
  const double span = 2.0 * (1.0 + 0.44) * std::max(c->rc_short, c->rc_long);
  std::vector<unsigned char> open(n, 0), ap(nsys, 0);
  for (int s = 0; s < nsys; s++)
    for (int i = c->off_h[s]; i < c->off_h[s + 1]; i++)
      for (int d = 0; d < 3; d++) {
        const double L0 = c->Lh[3 * s + d], v = pbc[3 * i + d];
        if (v == L0) continue;
        if (v != kNoPbc) {
          c->err = "mdg_set_atom_pbc: an atom's PBC on each axis is its system's box value or NO_PBC (atom " +
                   std::to_string(i) + ")";
          return 1;
        }
        if (L0 < span) {
          c->err = "mdg_set_atom_pbc: a box with atoms open on its axis " + std::to_string(d) +
                   " must be at least 2 (1 + 0.44) Rc = " + std::to_string(span / kAo) + " A wide [system " +
                   std::to_string(s) + "]";
          return 1;
        }
        open[i] |= (unsigned char)(1u << d);
        ap[s] = 1;
      }
  // each atom wrapped by its own PBC (Atom::applyPBC, Atom.cxx:93-125)
  std::vector<double> q(xyz, xyz + 3 * (size_t)n);
  std::vector<int32_t> pc(3 * (size_t)n, 0);
  if (pbc_count) pc.assign(pbc_count, pbc_count + 3 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      const double L = pbc[3 * i + d];
      double& x = q[3 * i + d];
      int32_t& k = pc[3 * i + d];
      if (L == kNoPbc) {
        if (k != 0) {  // Atom.cxx:105,114,123
          c->err = "mdg_set_atom_pbc: nonzero PBC_count on an open axis (Atom.cxx:105) [atom " + std::to_string(i) +
                   "]";
          return 1;
        }
        continue;
      }
      int guard = 0;
      while (x < 0 && guard++ < 1000000) { x += L; --k; }
      while (x >= L && guard++ < 1000000) { x -= L; ++k; }
    }
  bool any = false;
  for (unsigned char m : ap) any = any || m;
  c->apbc_h = ap;
  c->open_h = open;
  for (int s = 0; s < nsys; s++)
    if (bath_pbc_check(c, s, c->bath_h[s].geom)) return 1;
  if (!c->apbc_d && (dalloc(c, c->apbc_d, nsys) || dalloc(c, c->open_d, n))) return 1;
  HIPCHK(c, hipMemcpy(c->apbc_d, ap.data(), nsys, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->open_d, open.data(), n, hipMemcpyHostToDevice));
  c->box.apbc = any ? c->apbc_d : nullptr;
  c->box.open_axes = c->open_d;
  c->box.pcx = c->A.pcx;
  c->box.pcy = c->A.pcy;
  c->box.pcz = c->A.pcz;
  c->A.open_axes = any ? c->open_d : nullptr;
  c->A.boxL = c->box.L;
  if (setup_grids(c, q)) return 1;
  drop_graphs(c);  // the kernels' Box and Atoms arguments changed
  return mdg_set_positions(c, q.data(), pc.data(), nullptr);  // every list is rebuilt
}

int mdg_neighbor_counts(mdg_ctx* c, int32_t* counts) {
  int n = c->n;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int> s(n, 0), l(n, 0), ty(n), el(n);
  if (c->have_short) HIPCHK(c, hipMemcpy(s.data(), c->A.nls_cnt, n * 4, hipMemcpyDeviceToHost));
  if (c->have_long) HIPCHK(c, hipMemcpy(l.data(), c->A.nll_cnt, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ty.data(), c->A.type, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(el.data(), c->A.elem, n * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    bool long_handled = (c->slot_fbzl >= 0) || (c->slot_lj >= 0 && ty[i] <= T_H) ||
                        (c->slot_flj >= 0 && ty[i] != T_ION) ||  // FLJ handles H, C, Cu, Au
                        (c->slot_eam >= 0 && ty[i] == T_OTHER) ||  // Ackland handles Cu, Ag, Au
                        (c->slot_fbm >= 0 && (el[i] == 64 || el[i] == 40)) || (c->slot_tb >= 0 && el[i] == 64);
    counts[i] = s[i] + (long_handled ? l[i] : 0);
  }
  return 0;
}

int mdg_profile(mdg_ctx* c, int on) {
  if (on) {
    collect_profile(c);
    c->kst.clear();
  }
  c->profiling = on != 0;
  return 0;
}

int mdg_kernel_time(mdg_ctx* c, const char* kernel, double* total_ms, int64_t* launches) {
  collect_profile(c);
  auto it = c->kst.find(kernel);
  if (it == c->kst.end()) {
    *total_ms = 0;
    *launches = 0;
    return 0;
  }
  *total_ms = it->second.ms;
  *launches = it->second.launches;
  return 0;
}

int mdg_autotune(mdg_ctx* c, int reps, double* report) {
  if (c->n == 0) {
    c->err = "no system";
    return 1;
  }
  if (reps < 1) reps = 1;
  bool was = c->profiling;
  c->profiling = true;
  collect_profile(c);
  c->kst.clear();
  // Force calls leave the trajectory's state untouched (positions, velocities,
  // lists between rebuilds), so candidates are timed on repeated calls.
  auto measure = [&](const char* kernel) -> double {
    c->kst.clear();
    for (int r = 0; r < reps; r++) force_call(c);
    collect_profile(c);
    if (!kernel) {  // the four scattering force kernels together
      double t = 0.0;
      for (const char* k : {"k_lj_wave", "k_rebo_bonds", "k_bstar", "k_etors"}) {
        auto it = c->kst.find(k);
        if (it != c->kst.end()) t += it->second.ms / reps;
      }
      return t;
    }
    auto it = c->kst.find(kernel);
    return it == c->kst.end() ? 0.0 : it->second.ms / reps;
  };
  force_call(c);  // lists, bond array and queue sizes for the candidates
  if (check_err(c)) {
    c->profiling = was;
    return 1;
  }
  int nb = 0, nq = 0;
  HIPCHK(c, hipMemcpy(&nb, c->nbonds, 4, hipMemcpyDeviceToHost));
  HIPCHK(c, queued(c, nq));
  auto grids = [](long long units, int per_block) {
    std::vector<int> g;
    for (int v : {1024, 2048, 4096, 8192, 16384, 32768}) g.push_back(v);
    long long one = (units + per_block - 1) / per_block;
    if (one >= 1 && one < (1 << 20)) g.push_back((int)one);
    return g;
  };
  std::map<std::string, double> base_ms, best_ms;  // per kernel: untuned, tuned
  auto tune_knob = [&](const char* kernel, int& knob, const std::vector<int>& cand) {
    double t0 = measure(kernel);
    if (!kernel) {  // a knob of all scattering kernels: compared on their sum only
      int best = knob;
      double bt = t0;
      for (int v : cand) {
        int keep = knob;
        knob = v;
        double t = measure(kernel);
        if (t < bt) {
          bt = t;
          best = v;
        }
        knob = keep;
      }
      knob = best;
      return;
    }
    if (!base_ms.count(kernel)) base_ms[kernel] = t0;
    int best = knob;
    double bt = t0;
    for (int v : cand) {
      int keep = knob;
      knob = v;
      double t = measure(kernel);
      if (t < bt) {
        bt = t;
        best = v;
      }
      knob = keep;
    }
    knob = best;
    best_ms[kernel] = bt;
  };
  Tune& tu = c->tune;
  tune_knob(nullptr, c->A.wg_map, {0, 1});
  if (c->slot_rebo >= 0) {
    if (c->mixed) tune_knob("k_rebo_bonds", tu.occ_bonds, {1, 2, 3, 4, 11, 12});
    else tune_knob("k_rebo_bonds", tu.occ_bonds, {1, 2, 11});
    tune_knob("k_rebo_bonds", tu.grid_bonds, grids(nb, 256));
  }
  if (c->slot_lj >= 0) {
    if (c->mixed) tune_knob("k_lj_wave", tu.lj_shape, {0, 1, 2, 3, 4, 5});
    else tune_knob("k_lj_wave", tu.lj_shape, {0, 1, 2, 3});
    int lw = (tu.lj_shape >= 3) ? 16 : (tu.lj_shape == 2) ? 32 : 64;
    tune_knob("k_lj_wave", tu.grid_lj, grids(c->n, 4 * (64 / lw)));
    if (c->mixed) tune_knob("k_bstar", tu.occ_bstar, {1, 2, 3, 4});
    else tune_knob("k_bstar", tu.occ_bstar, {1, 2});
    tune_knob("k_bstar", tu.grid_bstar, grids(nq, 256));
  }
  if (c->slot_tors >= 0) tune_knob("k_etors", tu.grid_tors, grids(nb, 256));
  c->kst.clear();
  c->profiling = was;
  for (auto& e : c->graph_exec)  // captured launch shapes are stale now
    if (e) {
      hipGraphExecDestroy(e);
      e = nullptr;
    }
  double before = 0, after = 0;
  for (auto& kv : base_ms) before += kv.second;
  for (auto& kv : best_ms) after += kv.second;
  if (report) {
    double r[10] = {(double)tu.occ_bonds, (double)tu.grid_bonds, (double)tu.occ_bstar, (double)tu.grid_bstar,
                    (double)tu.grid_lj,   (double)tu.grid_tors,  before,                after,
                    (double)tu.lj_shape,  (double)c->A.wg_map};
    for (int q = 0; q < 10; q++) report[q] = r[q];
  }
  return check_err(c);
}

int mdg_set_graph_block(mdg_ctx* c, int steps) {
  if (steps < 1 || steps > 4096) {
    c->err = "mdg_set_graph_block: steps per launch must be in [1, 4096]";
    return 1;
  }
  if (steps != c->graph_block) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 2; k < 4; k++)
      if (c->graph_exec[k]) {
        hipGraphExecDestroy(c->graph_exec[k]);
        c->graph_exec[k] = nullptr;
      }
  }
  c->graph_block = steps;
  return 0;
}

int mdg_set_tune(mdg_ctx* c, const double* knobs) {
  const int occ_bonds = (int)knobs[0], occ_bstar = (int)knobs[2], lj_shape = (int)knobs[8], xcd = (int)knobs[9];
  const int max_occ = c->mixed ? 4 : 2;
  const bool shell = occ_bonds == 11 || (c->mixed && occ_bonds == 12);
  if (((occ_bonds < 1 || occ_bonds > max_occ) && !shell) || occ_bstar < 1 || occ_bstar > max_occ || lj_shape < 0 ||
      lj_shape > (c->mixed ? 5 : 3) || (xcd != 0 && xcd != 1)) {
    c->err = "mdg_set_tune: a knob outside the autotuner's candidates";
    return 1;
  }
  for (int q : {1, 3, 4, 5})
    if (knobs[q] < 0 || knobs[q] > (1 << 20)) {
      c->err = "mdg_set_tune: grid size outside [0, 2^20]";
      return 1;
    }
  Tune& tu = c->tune;
  tu.occ_bonds = occ_bonds;
  tu.grid_bonds = (int)knobs[1];
  tu.occ_bstar = occ_bstar;
  tu.grid_bstar = (int)knobs[3];
  tu.grid_lj = (int)knobs[4];
  tu.grid_tors = (int)knobs[5];
  tu.lj_shape = lj_shape;
  c->A.wg_map = xcd;
  drop_graphs(c);  // captured launch shapes are stale now
  return 0;
}

int mdg_event_record(mdg_ctx* c, int slot) {
  if (slot < 0 || slot >= 8) {
    c->err = "event slot out of range (0-7)";
    return 1;
  }
  if (!c->timer_ev[slot]) HIPCHK(c, hipEventCreate(&c->timer_ev[slot]));
  HIPCHK(c, hipEventRecord(c->timer_ev[slot], c->stream));
  return 0;
}

int mdg_event_elapsed(mdg_ctx* c, int a, int b, double* ms) {
  if (a < 0 || a >= 8 || b < 0 || b >= 8 || !c->timer_ev[a] || !c->timer_ev[b]) {
    c->err = "event not recorded";
    return 1;
  }
  HIPCHK(c, hipEventSynchronize(c->timer_ev[b]));
  float v = 0;
  HIPCHK(c, hipEventElapsedTime(&v, c->timer_ev[a], c->timer_ev[b]));
  *ms = v;
  return 0;
}

int mdg_synchronize(mdg_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mdg_bstar_bounds(mdg_ctx* c, double out[4]) {
  out[0] = c->host_T.glb[0];
  out[1] = c->host_T.glb[1];
  out[2] = c->host_T.plb[T_C];
  out[3] = c->host_T.plb[T_H];
  return 0;
}

int mdg_work_counts(mdg_ctx* c, int64_t out[3]) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<long long> t((size_t)kTallySlots * kTallyStride);
  HIPCHK(c, hipMemcpy(t.data(), c->tally, t.size() * sizeof(long long), hipMemcpyDeviceToHost));
  for (int k = 0; k < 3; k++) {
    long long v = 0;
    for (int q = 0; q < kTallySlots; q++) v += t[(size_t)q * kTallyStride + 3 + k];
    out[k] = v;
  }
  return 0;
}

int mdg_counters(mdg_ctx* c, int64_t out[8]) {
  long long h[8];
  int nb = 0, nq = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(h, c->counters, sizeof(h), hipMemcpyDeviceToHost));
  std::vector<long long> t((size_t)kTallySlots * kTallyStride);
  HIPCHK(c, hipMemcpy(t.data(), c->tally, t.size() * sizeof(long long), hipMemcpyDeviceToHost));
  long long tk[3] = {0, 0, 0};
  for (int q = 0; q < kTallySlots; q++)
    for (int k = 0; k < 3; k++) tk[k] += t[(size_t)q * kTallyStride + k];
  HIPCHK(c, hipMemcpy(&nb, c->nbonds, 4, hipMemcpyDeviceToHost));
  HIPCHK(c, queued(c, nq));
  int n = c->n;
  std::vector<int> cs(n), cl(n);
  HIPCHK(c, hipMemcpy(cs.data(), c->A.nls_cnt, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(cl.data(), c->A.nll_cnt, n * 4, hipMemcpyDeviceToHost));
  long long ss = 0, sl = 0;
  for (int i = 0; i < n; i++) {
    ss += cs[i];
    sl += cl[i];
  }
  out[0] = nb;
  out[1] = tk[0];
  out[2] = nq;
  out[3] = tk[1];
  out[4] = ss;
  out[5] = sl;
  out[6] = h[0] + h[1];
  out[7] = tk[2];
  return 0;
}
}
