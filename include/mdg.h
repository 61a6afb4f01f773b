/* mdg.h - C ABI of the MI355X force-and-integration engine for MDTK-style
 * hydrocarbon molecular dynamics (REBO/Brenner, AIREBO-LJ, torsion, ZBL and
 * the velocity-Verlet step).
 *
 * All arguments are plain pointers and sizes; no C++ or torch types cross the
 * boundary, no exceptions escape it.  Every call returns 0 on success and a
 * nonzero status otherwise; mdg_last_error() describes the failure.
 * Units are the reference's CGS units: positions in cm, velocities in cm/s,
 * energies in erg, gradients in erg/cm (mdtk/consts.cxx:28-38).  An open
 * (non-periodic) axis is given as NO_PBC = 1e-4 cm (mdtk/consts.hpp:96).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef MDG_H
#define MDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdg_ctx mdg_ctx;

/* Potential chain members, in the order applications/common.h:46-68 adds
 * them: FBZL (ZBL ion-target), FLJ, Brenner (replaces REBO), REBO,
 * AIREBO-LJ (requires REBO), ETors, Ackland; then FBM and TightBinding,
 * which the reference ships but does not register (TightBinding is
 * commented out at common.h:68-71; FBM is not named there). */
enum {
  MDG_POT_FBZL = 1,
  MDG_POT_BRENNER = 2,
  MDG_POT_REBO = 4,
  MDG_POT_AIREBO = 8,
  MDG_POT_ETORS = 16,
  MDG_POT_FLJ = 32,     /* FLJ (pairwise/FLJ.cxx): {H,C} x {Cu,Au}, chained after FBZL (common.h:49) */
  MDG_POT_ACKLAND = 64, /* Ackland EAM (manybody/Ackland/Ackland.cxx): Cu, Ag, Au; last in the chain (common.h:67); FP64 in both precisions */
  MDG_POT_FBM = 128,    /* FBM (pairwise/FBM.cxx): Born-Mayer Cu-Cu, Cu-Ar, cutoff 3.1-3.5 A; not in common.h's chain, follows it; FP64 */
  MDG_POT_TB = 256      /* TightBinding (manybody/TightBinding/TightBinding.cxx): Cu second moment; not in common.h's chain, follows FBM; FP64 */
};

enum { MDG_PREC_FP64 = 0, MDG_PREC_MIXED = 1 };

typedef struct mdg_config {
  int pot_mask;          /* OR of MDG_POT_* */
  int brenner_paramset;  /* 0: Brenner::POTENTIAL1, 1: POTENTIAL2 (Brenner.cxx:497-611) */
  int etors_compat;      /* 1: reference torsion derivative (AIREBO_ETors.cxx:110-111), 0: exact adjoint */
  int precision;         /* MDG_PREC_FP64 or MDG_PREC_MIXED */
  double fbzl_r0;        /* FBZL switch start/end, cm (common.h:46: 5.0/5.5 Angstrom) */
  double fbzl_r1;
  int max_short_nb;      /* capacity of one atom's REBO-range list (0: default 64) */
  int max_long_nb;       /* capacity of one atom's LJ-range list (0: default 512) */
} mdg_config;

/* Replaces constructing the potentials and FProxy::addPotential
 * (applications/common.h:36-68, mdtk/potentials/FProxy.hpp:37-40). */
int mdg_create(const mdg_config* cfg, mdg_ctx** out);

/* Replaces FProxy::freePotentials / ~SimLoop (FProxy.hpp:49-53, SimLoop.cxx:136-139). */
void mdg_destroy(mdg_ctx* ctx);

/* Message for the last failing call on this context (or on mdg_create). */
const char* mdg_last_error(const mdg_ctx* ctx);

/* Uploads the system: replaces AtomsArray::prepareForSimulatation,
 * checkMIC and FProxy::NL_init (AtomsContainer.cxx:106-123,
 * SimLoop.cxx:236-250).  elem: mdtk ElementID (H=1, C=12, Ar=40, Xe=131);
 * xyz/vel: n*3 interleaved; pbc[3]: box lengths, NO_PBC for an open axis;
 * fixed: n flags (may be null).  Positions are wrapped into the box. */
int mdg_set_system(mdg_ctx* ctx, int n, const int32_t* elem, const double* xyz, const double* vel,
                   const double pbc[3], const uint8_t* fixed);

/* A batch of nsys independent systems in one context, advanced together by
 * every call (one launch per kernel for the whole batch): the reference runs
 * one SimLoop per trajectory (e.g. one per bombardment in
 * applications/mdtrajsim/main.cpp:280-348, SimLoop.cxx:252-445); this replaces nsys of
 * them.  counts[nsys] atoms per system; elem/xyz/vel/fixed hold the systems'
 * atoms one after another (sum(counts) entries, as mdg_set_system); pbc[3*nsys]
 * each system's box.  Systems never interact; each keeps its own neighbour
 * lists, skin test and adaptive dt.  Calls that return per-system values
 * (mdg_energy_grad's e_terms, mdg_step_nve's outputs) then return nsys of
 * them, system-minor.  No single reference interface: the batch is the
 * caller's loop over trajectories. */
int mdg_set_systems(mdg_ctx* ctx, int nsys, const int32_t* counts, const int32_t* elem, const double* xyz,
                    const double* vel, const double* pbc, const uint8_t* fixed);

/* Number of systems in the context (1 after mdg_set_system). */
int mdg_num_systems(const mdg_ctx* ctx);

/* New positions for the host-driven loop, where the caller's SimLoop moves
 * the atoms (SimLoop.cxx:323-332): xyz[n*3] (cm, wrapped into the box as
 * Atom::applyPBC leaves them), pbc_count[n*3] (may be null: unchanged) and
 * dr[n*3], each atom's displacement since the previous call as
 * FGeneral::incDisplacement receives it (FGeneral.hpp:87-91).  dr feeds the
 * skin test of the next force call (NeighbourList.cxx:108-133); dr == null
 * forces a rebuild of every neighbour list. */
int mdg_set_positions(mdg_ctx* ctx, const double* xyz, const int32_t* pbc_count, const double* dr);

/* Each atom's own PBC vector (Atom::PBC, mdtk/Atom.hpp:35-82), for systems
 * whose atoms do not all share their box: pbc[n*3] (cm, NO_PBC for an open
 * axis), with positions xyz[n*3] and image counters pbc_count[n*3] (may be
 * null: zero) as the caller's atoms hold them; each atom is wrapped by its
 * own PBC (Atom::applyPBC, Atom.cxx:93-125; a nonzero counter on an open
 * axis is an error, Atom.cxx:105).  Every pair then takes depos with both
 * atoms' PBC (Atom.hpp:115-140): an open atom sees a periodic one at its
 * unwrapped coordinate, x + PBC*PBC_count.  On each axis an atom holds its
 * system's box value (mdg_set_system's pbc) or NO_PBC, as a SimLoopSaver
 * state's per-atom enabled flags express it, and a box with open atoms on
 * an axis must be at least 2 (1 + 0.44) Rc wide there (errors otherwise).
 * Such a system's neighbour lists test all its pairs, as the reference's do.
 * Call after mdg_set_system(s); every list is rebuilt. */
int mdg_set_atom_pbc(mdg_ctx* ctx, const double* pbc, const double* xyz, const int32_t* pbc_count);

/* Per-atom neighbour counts over all chain members' lists, as
 * FProxy::hasNeighbors tests them (FProxy.hpp:99-104): counts[n] > 0 iff
 * the atom has a neighbour in some list. */
int mdg_neighbor_counts(mdg_ctx* ctx, int32_t* counts);

/* Number of chain members and the MDG_POT_* kind of member i. */
int mdg_num_terms(const mdg_ctx* ctx);
int mdg_term_kind(const mdg_ctx* ctx, int i);

/* One force call at the current positions: replaces FProxy::operator()
 * (FProxy.cxx:45-61) with each FGeneral::operator() (FGeneral.hpp:94).
 * e_terms[nsys*mdg_num_terms]: energy of each member (erg), system s at
 * e_terms[s*mdg_num_terms + t]; grad[n*3]: dE/dx
 * (erg/cm), i.e. Atom::grad after the call.  Rebuilds a neighbour list first
 * when its skin test asks for it (NeighbourList.cxx:108-133). */
int mdg_energy_grad(mdg_ctx* ctx, double* e_terms, double* grad);

/* mdg_create + mdg_set_system + mdg_energy_grad + mdg_destroy in one call
 * that allocates no device memory, for callers that cannot allocate: tight
 * serving loops, harnesses that forbid hipMalloc.  No counterpart in the
 * reference.  Same arguments and results as those calls (host arrays: elem[n],
 * xyz[n*3] in cm, pbc[3], e_terms[mdg_num_terms], grad[n*3]); every device
 * buffer of the call comes from one static 64 MiB array per GPU, held until
 * the process ends, so n is limited to 4096 atoms (and the system's cell grid
 * and lists must fit; a larger one fails with a message).  The first call on a
 * device looks up that array's address.  Works on the null stream.  Calls
 * that overlap on one GPU (from two threads) share the array and race.
 * Errors: nonzero return, message from mdg_last_error(NULL). */
int mdg_energy_grad_noalloc(const mdg_config* cfg, int n, const int32_t* elem, const double* xyz,
                            const double pbc[3], double* e_terms, double* grad);

/* Differentiable parameters theta[MDG_NUM_PARAMS] (CGS): for the type pairs
 * p = 0 (C-C), 1 (C-H), 2 (H-H), REBO's pair parameters {Q, A, alpha, B1, B2,
 * B3, beta1, beta2, beta3} at theta[9p..9p+8] (REBO::REBO, REBO.cxx:856-941;
 * protected members REBO.hpp:155-175) and the LJ {epsilon, sigma} at
 * theta[27+2p], theta[28+2p] (AIREBO::setupPotential, AIREBO_LJ.cxx:229-247;
 * sigma also sets the S_r range [sigma, 2^(1/6) sigma], :239-247).  The
 * reference sets them only in its constructors; mdg_set_params replaces a
 * subclass that overwrites them.  mdg_energy_grad_params is mdg_energy_grad
 * plus dE_s/dtheta of every system s (dtheta[nsys*MDG_NUM_PARAMS], system s at
 * dtheta[s*MDG_NUM_PARAMS]), from adjoint kernels of the FP64 form; no
 * reference counterpart (the reference has no parameter derivatives). */
enum { MDG_NUM_PARAMS = 33 };
int mdg_get_params(const mdg_ctx* ctx, double* theta);
int mdg_set_params(mdg_ctx* ctx, const double* theta);
int mdg_energy_grad_params(mdg_ctx* ctx, double* e_terms, double* grad, double* dtheta);

/* Virial (no reference counterpart: MDTK computes none).  With
 * mdg_set_virial(ctx, 1) every force call also reduces, inside the force
 * kernels and in FP64, the per-system virial W = sum over the terms' pair
 * pushes of r_ab (x) F_ab (each force term is a sum of equal and opposite
 * pushes along its own minimum-image pair vectors), i.e. W = sum_i x_i (x) F_i
 * for an open system; six components xx, yy, zz, xy, xz, yz in erg, system s
 * at [6 s].  mdg_virial returns that of the last force call; mdg_step_virial
 * that of every step of the last mdg_step_nve (W at x(t+dt), beside epot),
 * step k of system s at [(k nsys + s) * 6].  Parameter-gradient calls
 * (mdg_energy_grad_params) leave it out.  Call after mdg_set_system(s). */
int mdg_set_virial(mdg_ctx* ctx, int on);
int mdg_virial(mdg_ctx* ctx, double* out);
int mdg_step_virial(mdg_ctx* ctx, double* out);

/* Neighbour list of chain member `term` as the reference would hold it
 * (FGeneral::nl, NeighbourList.hpp:41-83): counts[n], then the lists
 * concatenated (ascending index per atom) into idx[cap]. */
int mdg_nl_export(mdg_ctx* ctx, int term, int32_t* counts, int32_t* idx, int64_t cap);

/* nsteps iterations of SimLoop::executeMain's loop (SimLoop.cxx:252-445):
 * eq.1 drift, skin test and rebuild, force call, eq.2-4 kick (with the
 * thermal bath of mdg_set_thermal_bath, NVE without one), PBC wrap, dt
 * rule.  fixed_dt > 0 fixes dt (as a doBeforeIteration hook would);
 * otherwise dt = 0.05 A / v_max clamped to [1e-20, 5e-16] s (:409-416).
 * Per step outputs (each may be null): epot = E_pot(t+dt) (erg), ekin =
 * E_kin after the velocity update (erg), dt_used (s), netforce = |sum F|;
 * each holds nsteps*nsys values, step k of system s at [k*nsys + s]. */
int mdg_step_nve(mdg_ctx* ctx, int nsteps, double fixed_dt, double* epot, double* ekin, double* dt_used,
                 double* netforce);

/* Device errors of one system (a capacity such as 16 bonds or 512 long-list
 * neighbours of one atom, or a range check such as cos(theta) beyond +-1, the
 * reference's REQUIREs) stop that system only: inside mdg_step_nve it stands
 * still from the step after the error (its simTimeFinal is set to its
 * simTime, as if its run had ended), the other systems run every step, every
 * output is filled, and the call returns MDG_SYSTEMS_STOPPED with the
 * systems and errors in mdg_last_error.  Errors of shared resources (the b*
 * work queue) fail the call.  mdg_system_errors returns each system's error
 * bits (MDG_ERR_*; 0 = none) seen since the last call and clears them; a
 * stopped system stays stopped until its bits are cleared and mdg_set_time
 * gives it a new final time.  No reference counterpart: the reference runs
 * one trajectory, which a REQUIRE ends (SimLoop.cxx:176-205). */
enum { MDG_SYSTEMS_STOPPED = 2 };
enum {
  MDG_ERR_SHORT_NL_OVERFLOW = 1,  /* an atom's short-list neighbours beyond max_short_nb */
  MDG_ERR_LONG_NL_OVERFLOW = 2,   /* an atom's long-list neighbours beyond max_long_nb */
  MDG_ERR_BOND_OVERFLOW = 4,      /* more than 16 bonds (inside the REBO/Brenner f_c range) on one atom */
  MDG_ERR_COS_RANGE = 8,          /* cos(theta) outside [-1-1e-7, 1+1e-7] (FGeneral.hpp:147) */
  MDG_ERR_NEG_COORD = 16,         /* coordination below -1e-15 (REBO.cxx:434) */
  MDG_ERR_QUEUE_OVERFLOW = 32,    /* a shared work queue: fails the call */
  MDG_ERR_SPLINE_DOMAIN = 64,     /* a spline argument outside its domain */
  MDG_ERR_EAM_ROW_OVERFLOW = 128, /* more than 128 metal atoms inside the Ackland range of one atom */
  MDG_ERR_PATH_OVERFLOW = 256,    /* an atom's LJ pairs or C_ij paths beyond their capacity */
  MDG_ERR_TYPE_RANGE = 512        /* a packed position carries a type outside the tables */
};
int mdg_system_errors(mdg_ctx* ctx, uint32_t* out);

/* Thermal bath of one system (SimLoop::thermalBathCommon / thermalBathGeom*,
 * SimLoop.hpp:118-189), applied inside mdg_step_nve's kick as the reference
 * applies it (SimLoop.cxx:280, 352-386): atoms (non-fixed, opted in, inside
 * the geometry at x(t+dt)) get the force -V M gamma (1 - sqrt(To/T)), T the
 * kinetic temperature of the atoms inside at the start of the step (To/T
 * clamped to [-5, 5], 0 when |T| < 1e-5).  Geometries: universe (all atoms),
 * box (z > z_min, or within d_boundary of a lateral face and z > z_min_free;
 * needs x/y PBC), sphere (outside radius around center and z > z_min_free;
 * needs open boundaries).  Lengths in cm, To in K, gamma in 1/s; defaults of
 * the reference: gamma 1e13, z_min 1e6 A, d_boundary 0, z_min_free -5 A,
 * radius 1e6 A.  system -1 sets every system.  The energy each system took
 * from its bath since mdg_set_systems (SimLoop::check.energyTransferredFromBath,
 * :384) comes from mdg_bath_energy (out[nsys], erg); mdg_set_bath_atoms sets
 * Atom::apply_ThermalBath per atom (Atom.hpp:52; null = all). */
enum { MDG_TB_NONE = 0, MDG_TB_UNIVERSE = 1, MDG_TB_BOX = 2, MDG_TB_SPHERE = 3 };
typedef struct mdg_thermal_bath {
  int geometry;      /* MDG_TB_* */
  double To;         /* K */
  double gamma;      /* 1/s */
  double z_min;      /* box */
  double d_boundary; /* box */
  double z_min_free; /* box and sphere: zMinOfFreeZone */
  double center[3];  /* sphere */
  double radius;     /* sphere */
} mdg_thermal_bath;
int mdg_set_thermal_bath(mdg_ctx* ctx, int system, const mdg_thermal_bath* tb);
int mdg_set_bath_atoms(mdg_ctx* ctx, const uint8_t* apply);
int mdg_bath_energy(mdg_ctx* ctx, double* out);

/* SimLoop's clock per system (SimLoop.hpp fields simTime, simTimeFinal, dt;
 * loaded from a state's .time file by SimLoopSaver::load): mdg_set_time
 * sets them (system -1: every system; dt > 0); a system whose simTime has
 * reached simTimeFinal stands still in mdg_step_nve -- the loop condition
 * while (simTime < simTimeFinal) of SimLoop.cxx:252 -- and reports dt_used
 * 0, so the systems of a batch stop on their own clocks.  Each step adds its
 * dt to simTime (:405).  After mdg_set_systems: simTime 0, no final time,
 * dt 1e-20 s (SimLoop.cxx:46-50).  mdg_get_time returns per system
 * {simTime, simTimeFinal, next dt} (3*nsys values). */
int mdg_set_time(mdg_ctx* ctx, int system, double sim_time, double sim_time_final, double dt);
int mdg_get_time(mdg_ctx* ctx, double* out);

/* Trajectory output captured on the GPU inside mdg_step_nve, drained to host
 * memory after every chunk of steps (replaces SnapshotList::getSnapshot,
 * mdtrajsim/main.cpp:58-83 and SnapshotList.hpp:101-131, and
 * SimLoop::writetrajXVA, SimLoop.cxx:271-277,821-832).  At the start of
 * every iteration of a running system: a MDG_FRAME_SNAPSHOT frame with the
 * atoms flagged in snapshot_atoms (the reference selects those tagged
 * projectile or cluster; null = none) when CustomSimLoop's cadence is due
 * (t == 0 or int(t/I) != int((t - dt_prev)/I), I = 1e-14 s below 2 ps,
 * 5e-14 s below 4 ps, 1e-13 s beyond), and, with xva_frames != 0, a
 * MDG_FRAME_XVA frame with every atom of the system at iteration 0 and every
 * 1 ps below 2 ps, 5 ps beyond.  Records hold FP64 positions (cm),
 * velocities (cm/s), PBC counters and the atom's index in its system; the
 * SnapshotList keeps them as float32 (SnapshotList.hpp:38-48).
 * mdg_trajectory_size gives the frames and records gathered so far;
 * mdg_take_trajectory copies them out (frames[k].first indexes records) and
 * clears them.  Output stays on until mdg_set_trajectory_output(ctx, NULL, 0);
 * every call replaces the previous buffers and drops what they held. */
enum { MDG_FRAME_SNAPSHOT = 1, MDG_FRAME_XVA = 2 };
typedef struct mdg_frame {
  int32_t kind;      /* MDG_FRAME_* */
  int32_t system;
  int64_t iteration; /* SimLoop::iteration of the system */
  double sim_time;   /* s */
  double dt;         /* s, the dt of the iteration */
  int64_t first;     /* first record */
  int64_t count;     /* records */
} mdg_frame;
typedef struct mdg_atom_record {
  double x[3];
  double v[3];
  int32_t pbc_count[3];
  int32_t atom;      /* index within the system */
} mdg_atom_record;
int mdg_set_trajectory_output(mdg_ctx* ctx, const uint8_t* snapshot_atoms, int xva_frames);
int mdg_trajectory_size(mdg_ctx* ctx, int64_t* nframes, int64_t* nrecords);
int mdg_take_trajectory(mdg_ctx* ctx, mdg_frame* frames, mdg_atom_record* records);

/* Current state: replaces reading Atom::coords/V/an/PBC_count. */
int mdg_get_state(mdg_ctx* ctx, double* xyz, double* vel, double* acc, int32_t* pbc_count);

/* Accelerations acc[n*3] (cm/s^2) of loaded states, Atom::an as
 * SimLoopSaver::load reads it from <id>.a (SimLoopSaver.cxx:232-625), for the
 * systems with keep[s] = 1 (keep null: every system): the first
 * mdg_step_nve starts those systems from them instead of from the forces at
 * their positions, as SimLoop::execute does past iteration 0
 * (SimLoop.cxx:305-321 runs only at iteration 0), so that a run resumed from
 * a flushed state continues the interrupted one; the other systems start at
 * iteration 0 as usual.  Fixed atoms keep 0.  Only before the first
 * mdg_step_nve after mdg_set_systems. */
int mdg_set_accelerations(mdg_ctx* ctx, const double* acc, const uint8_t* keep);

/* SimLoop::iteration of one system or all (-1), as SimLoopSaver::loadIteration
 * sets it from the state id md%010d (SimLoopSaver.cxx:850-854); it counts
 * each system's own steps and names its XVA frames.  0 after
 * mdg_set_systems.  mdg_get_iteration returns nsys values. */
int mdg_set_iteration(mdg_ctx* ctx, int system, int64_t iteration);
int mdg_get_iteration(mdg_ctx* ctx, int64_t* out);

/* Runs the context's kernels on the given HIP stream (hipStream_t passed as
 * void*; null restores the context's own stream).  The stream must come from
 * the HIP runtime this library is linked against (a framework that bundles
 * its own HIP runtime has streams this library cannot use).  No reference
 * counterpart: lets a caller order the work on its own stream. */
int mdg_set_stream(mdg_ctx* ctx, void* stream);

/* Per-kernel timing with HIP events on the context's stream: on != 0 starts
 * collecting; mdg_kernel_time returns the summed milliseconds and launch
 * count of the named kernel since profiling was enabled.  No reference
 * counterpart (instrumentation). */
int mdg_profile(mdg_ctx* ctx, int on);
int mdg_kernel_time(mdg_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches);

/* Autotuner: times candidate launch shapes of the force kernels on the
 * context's own systems (reps force calls per candidate, which leave the
 * trajectory unchanged) and keeps the fastest for every later call: the
 * order in which the scattering kernels hand work to the XCDs (dealt in
 * turn, or a contiguous eighth of the work per XCD), the
 * waves-per-SIMD variant of the bond-order kernels (1-2; 1-4 in the mixed form; for the REBO bonds
 * also 11-12: LDS-staged shells, one wavefront per tile of 64 bonds, at 1 or 2 waves/SIMD; 12 in the
 * mixed form only), the shape of
 * the LJ kernel (shape 0-3: lanes per atom 64, 64 held to 3 waves/SIMD, 32,
 * 16; mixed form also 4-5: 16 lanes held to 4 or 3 waves/SIMD) and
 * the grid sizes of the bond, LJ, b* and torsion kernels.  report (may be
 * null) receives 10 values: occ_bonds, grid_bonds, occ_bstar, grid_bstar,
 * grid_lj, grid_tors (0 = built-in heuristic), the tuned kernels' summed ms
 * per call before and after, the LJ shape (0-5) and the XCD order (0: dealt
 * in turn, 1: contiguous eighths).  No reference
 * counterpart. */
int mdg_autotune(mdg_ctx* ctx, int reps, double* report);
/* Sets the launch shapes directly from a report of mdg_autotune (the same 10
 * values; the two timings are ignored), e.g. one saved from an earlier run
 * on the same workload.  A knob outside the autotuner's candidates is an
 * error.  No reference counterpart. */
int mdg_set_tune(mdg_ctx* ctx, const double* knobs);
/* Steps per graph launch in mdg_step_nve (default 32; 1..4096).  The whole
 * step of every system lives on the device (skin tests, rebuild flags,
 * clocks, per-step outputs), so a block of steps is captured into one
 * hipGraph and one launch advances the batch by that many steps; a chunk's
 * remainder runs one step per launch.  Results do not depend on it.  No
 * reference counterpart (SimLoop::executeMain's loop, SimLoop.cxx:252-445,
 * runs one iteration at a time). */
int mdg_set_graph_block(mdg_ctx* ctx, int steps);

/* Timing with HIP events on the context's stream: mdg_event_record records
 * event `slot` (0-7); mdg_event_elapsed waits for event b and returns the
 * milliseconds from a to b; mdg_synchronize waits for the stream.  No
 * reference counterpart (instrumentation). */
int mdg_event_record(mdg_ctx* ctx, int slot);
int mdg_event_elapsed(mdg_ctx* ctx, int a, int b, double* ms);
int mdg_synchronize(mdg_ctx* ctx);

/* Interaction counts of the last force call (for roofline accounting):
 * out[0] bonds, out[1] LJ pairs inside the cutoff, out[2] b* evaluations,
 * out[3] torsion quads, out[4] short-list entries, out[5] long-list entries,
 * out[6] neighbour-list rebuilds so far, out[7] LJ pair rows built so far
 * (one per atom and build; the rows are the long-list entries within the LJ
 * cutoff plus a 1 A margin).  No reference counterpart. */
int mdg_counters(mdg_ctx* ctx, int64_t out[8]);
/* Work of the last force call for the FLOP tally of the bond-order terms:
 * out[0] sigma-pi triplets of the REBO bonds (bond, side, neighbour k; per
 * centre with nb bonds nb (nb - 1)), out[1] sigma-pi triplets of the b*
 * bound pass (the neighbours of both atoms of every queued pair), out[2]
 * queued b* pairs whose bound failed (the full bond order and adjoint; the
 * mixed form's two-pass b*, else 0).  No reference counterpart. */
int mdg_work_counts(mdg_ctx* ctx, int64_t out[3]);
/* The table bounds behind the LJ pass's per-atom b* test: out[0] a lower
 * bound of G for a C centre (G_C1 and G_C2), out[1] of G_H, out[2] of P_CC,
 * out[3] of P_CH, each over its whole domain less a margin for the FP32
 * form.  No reference counterpart (diagnostics). */
int mdg_bstar_bounds(mdg_ctx* ctx, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
