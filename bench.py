"""Benchmark: batched AIREBO bombardment trajectories on one MI355X.

Headline workload (one "step" = one velocity-Verlet iteration of every
trajectory: drift, per-trajectory skin test and neighbour-list rebuild when
due, FBZL + REBO + AIREBO-LJ + torsion force call, kick, PBC wrap, per-
trajectory energies and adaptive-dt bookkeeping):
  256 independent polyethylene slabs, each ~U[2000, 12000] atoms (seeded,
  whole layers of 288 atoms; chains in the surface plane, x/y periodic, z
  open, bottom chain row fixed), 300 K Maxwell-Boltzmann velocities, one Ar
  ion 8 A above the surface at 100 eV, normal incidence, seeded impact point;
  dt = 0.1 fs fixed.  All trajectories are advanced by one hipGraph launch
  per 32 steps, in the mixed FP32/FP64 form.  An untimed approach (set-up,
  before the warm-up) brings the ions to 2 A above their surfaces, so the
  timed window holds the impacts and their list rebuilds.
The plain FP64 drop-in runs the same trajectories the way the reference
does, one trajectory per context and one after another; it is timed on a
sample of the trajectories.  Synthetic input; positions and velocities are
resident in HBM before every timed region.

Prints ONE JSON line.  `value` = sum over trajectories of atoms x steps /
wall seconds of the timed region (hipEvents on the engine's stream, bracketed
by device synchronisation).  `roofline` prices the dominant kernel against
MI355X's HBM peak by its algorithmic bytes; `cpu_baseline` times the oracle (the
CPU restatement of the reference's algorithm, single thread) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s on 1 GPU, AIREBO: mixed+batched vs plain FP64 at force err ≤1e-5"
POTS = ("FBZL", "REBO", "AIREBO", "ETors")

KERNELS = ("k_bstar", "k_rebo_bonds", "k_lj_wave", "k_pair_rows", "k_etors", "k_coord", "k_bond_enum", "k_nl_build_long",
           "k_nl_build_short", "k_kick", "k_drift", "k_sys_disp", "k_sys_step_end", "k_zero", "k_fbzl",
           "k_cell_assign", "k_cell_scan", "k_cell_fill", "k_cell_clear")
EAM_KERNELS = ("k_eam_density", "k_eam_force", "k_metal_rows", "k_nl_build_long", "k_kick", "k_drift")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ntraj", type=int, default=256)
    ap.add_argument("--dropin-sample", type=int, default=8, help="trajectories timed in the FP64 drop-in form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the single-system PE and drift figures")
    ap.add_argument("--approach", type=int, default=-1,
                    help="untimed steps before the warm-up (-1: until the ions are 2 A above their surfaces)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="profiling runs: only the mixed batched engine (tuned, timed, kernel times), no CPU leg")
    ap.add_argument("--workload", choices=("bombard", "metal"), default="bombard",
                    help="metal: only the Ackland Cu-crystal figure, as its own JSON line")
    return ap.parse_args()


def approach_steps(batch, dt, above=2.0):
    """Steps of dt until the ions, moving along -z at their initial speed,
    are `above` A over the top surface atom of their slab (the untimed
    approach, so that the timed window holds the impacts)."""
    import numpy as np
    n = []
    for s in batch:
        h = s.pos[-1, 2] - s.pos[:-1, 2].max()
        v = abs(s.vel[-1, 2]) * (dt / 1e-15)  # A/fs x fs per step
        n.append(max(0.0, (h - above) / v) if v > 0 else 0.0)
    return int(np.median(n))


def timed(eng, steps, warmup, dt, approach=0):
    """Milliseconds for `steps` Verlet iterations between two HIP events on the
    engine's stream, with the device synchronised before and after; the
    per-step outputs and the list rebuilds that fell inside the window.
    `approach` untimed steps run first (set-up), then `warmup`."""
    if approach:
        eng.run(approach, fixed_dt=dt)
    eng.run(max(warmup, 1), fixed_dt=dt)
    eng.synchronize()
    c0 = eng.counters()
    eng.record(0)
    res = eng.run(steps, fixed_dt=dt)
    eng.record(1)
    eng.synchronize()
    c1 = eng.counters()
    res["rebuilds"] = {"neighbour_lists (per system)": c1["rebuilds"] - c0["rebuilds"],
                       "lj_pair_rows (per atom)": c1["pair_rows_built"] - c0["pair_rows_built"]}
    return eng.elapsed_ms(0, 1), res


def kernel_times(eng, steps, dt):
    """Per-kernel HIP-event times over `steps` eager iterations (same stream)."""
    eng.profile(True)
    eng.run(steps, fixed_dt=dt)
    eng.synchronize()
    kt = {k: eng.kernel_time(k) for k in KERNELS}
    eng.profile(False)
    return kt


LJ_SHAPES = {0: (1, 64), 1: (3, 64), 2: (1, 32), 3: (1, 16), 4: (4, 16), 5: (3, 16)}  # autotuner's lj_shape -> (OCC, LW)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_bombard256.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_bombard256.json")
METAL_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_cu256k.json")


def lj_algorithmic_bytes(cnt, natoms):
    """Compulsory memory traffic of one k_lj_wave launch (every byte read or
    written once): one 4-byte index per LJ pair inside the cutoff (the half
    list, i < j); per atom its row count (4 B), packed position and type
    (32 B), bond count (4 B), gradient read + write (48 B) and energy (8 B);
    per bond, both directions' records (32 B) and partner index (4 B); per
    queued b* pair one 16-byte entry."""
    return (4.0 * cnt["lj_pairs"] + natoms * (4 + 32 + 4 + 48 + 8)
            + 2.0 * cnt["bonds"] * 36 + 16.0 * cnt["bstar"])


def measured_traffic(kernel, path=TRAFFIC_FILE):
    """Bytes beyond L2 per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    from the committed rocprofv3 counter summary of the workload."""
    try:
        with open(path) as f:
            t = json.load(f)["bytes_per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if t is None else t["read_bytes"] + t["write_bytes"]


def valu_profile(kernel, path=VALU_FILE):
    """FP32/FP64 FLOPs and VALU activity per launch of `kernel` from the
    committed rocprofv3 counter summary of the workload (or None)."""
    try:
        with open(path) as f:
            return json.load(f)["per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


# Algorithmic FLOPs per unit of work of each class of the AIREBO path, one
# forward and one adjoint pass as the kernels compute them (mdg_force.hip),
# tallied from the source with add/sub/mul/compare = 1, div = 8, sqrt = 8,
# exp/sin/cos = 20, pow = 40 (no FMA: the kernels build without contraction);
# integer and address work, loads and atomics are not counted, nor the
# reference's seeded re-evaluations.  Units and counts of the last force call
# come from mdg_counters / mdg_work_counts.
FLOPS_PER_UNIT = {
    # k_rebo_bonds outside the bond order: displacement and r in FP32 and
    # FP64 (35), VR with its derivative (55), the three VA exponentials (73),
    # b VA and its radial derivative (19), the gradient push (17); cutoffs
    # inside R1 are compares
    "rebo_pair (bond)": 204,
    # one (bond, side, neighbour k) of the sigma-pi sums: forward cos(theta)
    # (18) and G with both quintic pieces and the S'(N) switch (82), D += G f
    # (3); adjoint cos(theta) and G again (100), d cos / d(r_ij, r_ik) and the
    # three pushes (101), dG/dN and the scatter (13)
    "sigma_pi (triplet)": 318,
    # the rest of one bond order: six coordination exclusions (24), P bicubics
    # (4/3 per bond in PE, 140 each), D^-1/2 twice (80), the Nconj sums (26),
    # pi_rc tricubic (600), T_ij tricubic on the C-C third (200), the adjoint
    # seeds D^-3/2 (88) and the coordination and Nconj derivatives (~190)
    "bond_order_rest (bond)": 1300,
    # one ETors quad: the dihedral (two cross products, norms, the cosine and
    # sines, ~140), V(omega) and the cutoff weights (~40), the adjoint of the
    # dihedral into four atoms and the weight derivatives (~220)
    "etors (quad)": 400,
    # one LJ pair inside the cutoff: the pair-row sweep's distance tests
    # (15 per row entry, ~1.65 entries per pair, 25), displacement and r
    # (25), S_r, the cutoff (20 averaged over the 5.0-5.5 A shell),
    # (sigma/r)^6, ^12, VLJ and its derivative (40), the gradient push and
    # j-side scatter (20)
    "lj (pair)": 135,
    # one queued b* pair in the bound pass, besides its triplets: r, S_r,
    # C_ij weight, six coordination exclusions, the P bicubics (180), the
    # bound (37) and the LJ evaluation with S_r S_b and the C_ij adjoint (120)
    "bstar_bound (pair)": 400,
    # one b* triplet of the bound pass (forward sigma-pi only)
    "bstar_triplet (triplet)": 104,
    # one b* pair given the full bond order in the second pass: the bond
    # order's remainder and adjoint (1300) plus ~8 adjoint triplets (214 each)
    "bstar_full (pair)": 3000,
}


def flop_tally(cnt, work, step_ms):
    """Per-class algorithmic FLOPs of one force call (the last of the window)
    and the step's rate against the FP32 vector peak (the mixed form's
    bond-order, LJ and torsion algebra is FP32; its FP64 parts, the pair
    terms and the sums, are under a tenth of the tally)."""
    units = {"rebo_pair (bond)": cnt["bonds"], "sigma_pi (triplet)": work["rebo_triplets"],
             "bond_order_rest (bond)": cnt["bonds"], "etors (quad)": cnt["quads"], "lj (pair)": cnt["lj_pairs"],
             "bstar_bound (pair)": cnt["bstar"], "bstar_triplet (triplet)": work["bstar_triplets"],
             "bstar_full (pair)": work["bstar_full"]}
    rows = {k: {"units": int(units[k]), "flops_per_unit": f, "gflop": units[k] * f / 1e9}
            for k, f in FLOPS_PER_UNIT.items()}
    tot = sum(r["gflop"] for r in rows.values()) * 1e9
    rate = tot / (step_ms / 1e3) / 1e12
    return {"classes": rows, "gflop_per_step": tot / 1e9, "achieved": rate, "unit": "TFLOP/s",
            "peak_fp32": 157.3, "frac_fp32": rate / 157.3,
            "pi_dh": "not tallied: T_ij vanishes for the chains' saturated C-C bonds; the impacts' few "
                     "unsaturated ones are left out",
            "weights": "add/mul/compare 1, div 8, sqrt 8, exp/sin/cos 20, pow 40; forward + adjoint; "
                       "tallied from the kernel source"}


def roofline(kt, cnt, natoms, tune):
    """The dominant kernel against MI355X's HBM peak (8 TB/s): algorithmic
    bytes per launch over the launch's average HIP-event duration; beside it
    the same launch's VALU rate (FLOPs per launch from the committed counter
    summary over the live duration) against the FP32 and FP64 vector peaks."""
    dom = max(kt, key=lambda k: kt[k][0])
    dom_ms, dom_n = kt[dom]
    avg_s = dom_ms / max(dom_n, 1) / 1e3
    out = {"bound": "hbm", "kernel": dom, "achieved": None, "peak": 8000.0, "unit": "GB/s", "frac": None,
           "traffic": None, "avg_launch_ms": avg_s * 1e3, "launches": dom_n}
    if dom == "k_lj_wave":
        occ, lw = LJ_SHAPES[int(tune.get("lj_shape", 3))]
        name = f"k_lj_wave<float, false, {occ}, {lw}, false>"
        alg = lj_algorithmic_bytes(cnt, natoms)
        ach = alg / avg_s / 1e9
        out.update(kernel=name, achieved=ach, frac=ach / 8000.0, traffic=measured_traffic(name),
                   algorithmic_bytes=alg,
                   traffic_source=("profiles/traffic_bombard256.json: rocprofv3 FETCH_SIZE + WRITE_SIZE of this "
                                   "kernel on this workload, copy-calibrated (measured once, not in this run)"),
                   note="latency bound: waves wait on dependent list -> position -> bond-record gathers and on "
                        "the LDS gradient cache (see valu.busy)")
        v = valu_profile(name)
        lj_alg = cnt["lj_pairs"] * FLOPS_PER_UNIT["lj (pair)"] / avg_s / 1e12
        if v:
            f32 = v["fp32_flops"] / avg_s / 1e12
            f64 = v["fp64_flops"] / avg_s / 1e12
            out["valu"] = {"algorithmic": lj_alg, "algorithmic_frac_fp32": lj_alg / 157.3,
                           "algorithmic_source": "the LJ class of the FLOP tally (flops.classes) over this launch",
                           "achieved_fp32": f32, "peak_fp32": 157.3, "frac_fp32": f32 / 157.3,
                           "achieved_fp64": f64, "peak_fp64": 78.6, "frac_fp64": f64 / 78.6, "unit": "TFLOP/s",
                           "busy": v["valu_busy"], "lane_utilisation": v["lane_utilisation"],
                           "wait_fraction": v["wait_fraction"],
                           "source": "profiles/valu_bombard256.json (rocprofv3 SQ counters of this kernel)"}
            out["binding"] = "neither: latency (dependent gathers, LDS atomics)"
    return out


def eam_algorithmic_bytes(entries, inrange, natoms):
    """Compulsory traffic of the two Ackland passes.  Density pass: a 4-byte
    index per metal-row entry (partners within the range + 1 A), per atom the
    row count, element and packed
    position (4 + 1 + 32 B) read and F', F, the in-range count (16 + 8 + 4 B)
    written, and a 4-byte index per in-range partner written.  Force pass: a
    4-byte index per in-range partner, per atom the count, element, packed
    position and (F', element) (4 + 1 + 32 + 16 B), gradient read + write
    (48 B) and energy read + write (16 B)."""
    dens = 4.0 * entries + natoms * (37 + 28) + 4.0 * inrange
    force = 4.0 * inrange + natoms * (53 + 48 + 16)
    return dens, force


def inrange_pairs(pos_a, box_a, r_a):
    """Ordered pairs (i != j) closer than r_a in a periodic box."""
    from scipy.spatial import cKDTree
    import numpy as np
    p = np.mod(pos_a, box_a)
    t = cKDTree(p, boxsize=box_a)
    return float(t.count_neighbors(t, r_a) - len(p))


def metal_block(steps, warmup):
    """Ackland EAM alone on a periodic Cu crystal (40^3 cells, 256,000 atoms,
    300 K, dt 1 fs): atom-steps/s and both passes against the HBM peak."""
    import numpy as np
    from mdtk_amd import Engine
    from mdtk_amd import systems as S
    cu = S.fcc(64, 40, 40, 40)
    cu.vel = S.maxwell_velocities(cu, 300.0, stream=3)
    e = Engine.from_system(cu, potentials=("Ackland",))
    ms, res = timed(e, steps, warmup, 1e-15)
    e.profile(True)
    e.run(steps, fixed_dt=1e-15)
    e.synchronize()
    kt = {k: e.kernel_time(k) for k in EAM_KERNELS}
    e.profile(False)
    long_entries = float(e.counters()["long_entries"])
    x, _, _, _ = e.state()
    rng = 1.2247449 * 4.086  # Ackland's largest range (Ag), A
    inrange = inrange_pairs(x / S.ANG, cu.pbc, rng)
    entries = inrange_pairs(x / S.ANG, cu.pbc, rng + 1.0)  # the metal rows' radius (1 A skin)
    dens, force = eam_algorithmic_bytes(entries, inrange, cu.n)
    rl = {}
    for k, alg, full in (("k_eam_density", dens, "k_eam_density<mdg::AcklandModel>"),
                         ("k_eam_force", force, "k_eam_force<mdg::AcklandModel>")):
        t_ms, nl = kt[k]
        avg = t_ms / max(nl, 1)
        rl[k] = {"bound": "hbm", "algorithmic_bytes": alg, "avg_launch_ms": avg,
                 "achieved": alg / (avg / 1e3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                 "frac": alg / (avg / 1e3) / 1e9 / 8000.0, "traffic": measured_traffic(full, METAL_TRAFFIC_FILE)}
    et = (res["epot"] + res["ekin"]) / S.EV
    out = {"workload": f"Cu fcc 40x40x40 ({cu.n} atoms), periodic, 300 K, Ackland EAM alone, dt 1 fs, FP64",
           "value": cu.n * steps / (ms / 1e3), "unit": "atom-steps/s", "ms_per_step": ms / steps,
           "long_entries_per_atom": long_entries / cu.n, "metal_row_entries_per_atom": entries / cu.n,
           "inrange_partners_per_atom": inrange / cu.n,
           "kernels_ms_per_step": {k: v[0] / max(steps, 1) for k, v in kt.items()},
           "roofline": rl, "max_abs_dE_tot_eV": float(np.abs(et - et[0]).max())}
    return out


def force_rms(systems):
    """RMS over atoms of |F_mixed - F_fp64| in eV/A at the given configuration."""
    import numpy as np
    from mdtk_amd import Engine
    from mdtk_amd import systems as S
    e64 = Engine.from_systems(systems, potentials=POTS, etors_compat=False)
    _, g64 = e64.energy_grad()
    del e64
    emx = Engine.from_systems(systems, potentials=POTS, etors_compat=False, precision="mixed")
    _, gmx = emx.energy_grad()
    del emx
    return float(np.sqrt(((g64 - gmx) ** 2).sum(1).mean()) / (S.EV / S.ANG))


def nve_drift(sysm, precision, steps, dt):
    """Slope of E_tot/N (least squares) in eV/atom/ps, exact torsion adjoint."""
    import numpy as np
    from mdtk_amd import Engine
    from mdtk_amd import systems as S
    eng = Engine.from_system(sysm, potentials=("REBO", "AIREBO", "ETors"), etors_compat=False, precision=precision)
    r = eng.run(steps, fixed_dt=dt)
    e = (r["epot"] + r["ekin"]) / S.EV / sysm.n
    t = np.arange(steps) * dt * 1e12
    slope = float(np.polyfit(t, e, 1)[0])
    return slope, float(np.abs(e - e[0]).max())


def host_driven(traj, calls):
    """Force calls through host buffers, as the FGeneral adapter makes them
    (positions with a zero displacement in, energies and gradient out, PCIe
    both ways): atom-force-evaluations/s of one trajectory in FP64."""
    import numpy as np
    from mdtk_amd import Engine
    e = Engine.from_system(traj, potentials=POTS, etors_compat=False)
    x = e.state()[0]
    dr = np.zeros_like(x)
    e.set_positions(x, dr_cm=dr)
    e.energy_grad()
    t0 = time.perf_counter()
    for _ in range(calls):
        e.set_positions(x, dr_cm=dr)
        e.energy_grad()
    return traj.n * calls / (time.perf_counter() - t0)


def _cpu_worker(args):
    """One oracle process: a trajectory stepped for budget_s seconds."""
    traj, budget_s = args
    from oracle.oracle import Oracle, POT_FBZL, POT_REBO, POT_AIREBO, POT_ETORS
    o = Oracle.from_system(traj, pots=POT_FBZL | POT_REBO | POT_AIREBO | POT_ETORS, etors_compat=False)
    o.run(1, fixed_dt=0.1e-15)  # set-up and first step (list build, initial force) untimed
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.run(1, fixed_dt=0.1e-15)
        steps += 1
    return traj.n, steps, time.perf_counter() - t0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(batch, budget_s: float):
    """The oracle (single-threaded CPU restatement, as the reference runs one
    trajectory per process) on the batch's smallest trajectories, one process
    per host core (at most 16), all at once; run before the GPU is touched."""
    import multiprocessing as mp
    from oracle.oracle import build
    build()
    procs = max(1, min(16, os.cpu_count() or 1, len(batch)))
    sample = sorted(batch, key=lambda s: s.n)[:procs]
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_cpu_worker, [(t, budget_s) for t in sample])
    rate = sum(n * k / dt for n, k, dt in res)
    one = [n * k / dt for n, k, dt in res]
    return {"value": rate, "unit": "atom-steps/s", "cores": procs, "kind": "port", "cpu_model": cpu_model(),
            "per_process_value": sum(one) / len(one),
            "sample": f"the {procs} smallest bombardment trajectories of the batch "
                      f"({min(t.n for t in sample)}-{max(t.n for t in sample)} atoms: PE slab + Ar at 100 eV), "
                      f"FBZL+REBO+AIREBO-LJ+ETors FP64, dt=0.1 fs, one single-threaded oracle process each, "
                      f"{budget_s:.0f} s each, run together"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU per process; --gpus {a.gpus} is not supported "
              "(independent trajectories shard across GPUs with no communication).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from mdtk_amd import Engine
    from mdtk_amd import systems as S

    if a.workload == "metal":
        m = metal_block(a.steps, a.warmup)
        print(json.dumps({"metric": "atom-steps/s on 1 GPU, Ackland EAM (Cu crystal)", "n_gpus": 1, "steps": a.steps,
                          "warmup": a.warmup, "higher_is_better": True, "dtype": "f64", "data": "synthetic", **m}),
              flush=True)
        return

    dt = 0.1e-15
    batch = S.bombard_batch(a.ntraj)
    natoms = sum(s.n for s in batch)
    log(f"batch: {len(batch)} trajectories, {natoms} atoms")
    cpu = None
    if not a.no_cpu and not a.mixed_only:  # before any GPU call: the oracle processes are forked from this one
        cpu = cpu_baseline(batch, a.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.4e} atom-steps/s on {cpu['cores']} processes")

    napp = approach_steps(batch, dt) if a.approach < 0 else a.approach
    log(f"untimed approach: {napp} steps")

    # headline: mixed + batched, one graph launch per 32 steps for all trajectories
    emx = Engine.from_systems(batch, potentials=POTS, etors_compat=False, precision="mixed")
    tune_mx = emx.autotune(reps=2)
    log(f"mixed batched autotune: {tune_mx}")
    if not a.mixed_only and napp:  # the configuration the timed window starts from
        emx.run(napp, fixed_dt=dt)
        x0 = emx.state()[0]
        off = emx.offsets
        at_window = [S.System(s.elem, x0[off[k]:off[k + 1]] / S.ANG, s.pbc, None, s.fixed, s.name)
                     for k, s in enumerate(batch)]
        ms_mx, res_mx = timed(emx, a.steps, a.warmup, dt)
    else:
        at_window = batch
        ms_mx, res_mx = timed(emx, a.steps, a.warmup, dt, napp)
    v_mx = natoms * a.steps / (ms_mx / 1e3)
    log(f"mixed batched: {ms_mx / a.steps:.3f} ms/step, {v_mx:.4e} atom-steps/s")
    kt_mx = kernel_times(emx, a.steps, dt)
    cnt_mx = emx.counters()
    work_mx = emx.work_counts()
    flops_mx = flop_tally(cnt_mx, work_mx, ms_mx / a.steps)
    log(f"algorithmic FLOPs: {flops_mx['gflop_per_step']:.2f} GFLOP/step, {flops_mx['achieved']:.2f} TFLOP/s "
        f"({100 * flops_mx['frac_fp32']:.1f}% of the FP32 vector peak)")
    del emx
    if a.mixed_only:
        print(json.dumps({"metric": METRIC, "value": v_mx, "unit": "atom-steps/s", "n_gpus": 1, "steps": a.steps,
                          "warmup": a.warmup, "ms_per_step": ms_mx / a.steps, "mode": "mixed batched only",
                          "autotune": tune_mx, "roofline": roofline(kt_mx, cnt_mx, natoms, tune_mx),
                          "kernels_ms_per_step": {k: v[0] / max(a.steps, 1) for k, v in kt_mx.items()},
                          "counters": cnt_mx, "work": work_mx, "flops": flops_mx}), flush=True)
        return

    f_rms = force_rms(at_window)
    log(f"mixed force RMS error vs FP64 at the window's start: {f_rms:.3e} eV/A")

    # FP64 batched (same kernels in FP64)
    e64 = Engine.from_systems(batch, potentials=POTS, etors_compat=False)
    tune_64 = e64.autotune(reps=2)
    log(f"fp64 batched autotune: {tune_64}")
    ms_64b, _ = timed(e64, a.steps, a.warmup, dt, napp)
    v_64b = natoms * a.steps / (ms_64b / 1e3)
    log(f"fp64 batched: {ms_64b / a.steps:.3f} ms/step, {v_64b:.4e} atom-steps/s")
    del e64

    # plain FP64 drop-in: one trajectory per context, one after another
    sample = batch[:a.dropin_sample]
    tot_ms, tot_as = 0.0, 0
    for s in sample:
        e = Engine.from_system(s, potentials=POTS, etors_compat=False)
        ms, _ = timed(e, a.steps, a.warmup, dt, napp)
        tot_ms += ms
        tot_as += s.n * a.steps
        del e
    v_drop = tot_as / (tot_ms / 1e3)
    log(f"fp64 drop-in (sequential single trajectories): {v_drop:.4e} atom-steps/s")
    v_host = host_driven(sample[0], 50)
    log(f"fp64 force calls through host buffers: {v_host:.4e} atom-evaluations/s")

    use_mixed = f_rms <= 1e-5
    value = v_mx if use_mixed else v_64b
    ms = ms_mx if use_mixed else ms_64b
    rl = roofline(kt_mx, cnt_mx, natoms, tune_mx)
    et = (res_mx["epot"] + res_mx["ekin"]) / S.EV
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 mixed" if use_mixed else "f64",
        "data": "synthetic",
        "config": {"workload": (f"batched bombardment: {len(batch)} PE slabs ~U[2000,12000] atoms + Ar 100 eV, "
                                "FBZL+REBO+AIREBO-LJ+ETors, dt 0.1 fs, one hipGraph launch per 32 steps"),
                   "timed_window": (f"after an untimed approach of {napp} steps (the ions 2 A above their "
                                    f"surfaces) and {a.warmup} warm-up steps: the impacts"),
                   "trajectories": len(batch), "atoms": natoms,
                   "atoms_min": min(s.n for s in batch), "atoms_max": max(s.n for s in batch),
                   "precision": "mixed" if use_mixed else "fp64", "batched": True,
                   "torsion": "exact adjoint", "parallelism": "single GPU"},
        "mixed_batched_value": v_mx,
        "fp64_batched_value": v_64b,
        "fp64_dropin_value": v_drop,
        "fp64_dropin_sample": f"{len(sample)} trajectories, {sum(s.n for s in sample)} atoms, sequential contexts",
        "mixed_batched_over_fp64_dropin": v_mx / v_drop,
        "host_driven_fp64_value": v_host,
        "host_driven_sample": f"{sample[0].n}-atom trajectory, 50 force calls through host buffers "
                              "(positions in, energies and gradient out), FP64",
        "mixed_force_rms_err_eV_per_A": f_rms,
        "autotune": {"mixed_batched": tune_mx, "fp64_batched": tune_64},
        "roofline": rl,
        "kernels_ms_per_step": {k: v[0] / max(a.steps, 1) for k, v in kt_mx.items()},
        "counters": cnt_mx,
        "work": work_mx,
        "flops": flops_mx,
        "energy_check_eV": {"max_abs_dE_tot_over_steps": float(np.abs(et - et[0]).max())},
        "rebuilds_in_timed_window": res_mx["rebuilds"],
    }
    if not a.no_extra:
        pe = S.polyethylene(24, 24, 16)
        pe.vel = S.maxwell_velocities(pe, 300.0, stream=0)
        for prec in ("fp64", "mixed"):
            e = Engine.from_system(pe, potentials=("REBO", "AIREBO", "ETors"), precision=prec)
            ms_pe, _ = timed(e, a.steps, a.warmup, 0.2e-15)
            out[f"pe110k_{prec}_value"] = pe.n * a.steps / (ms_pe / 1e3)
            del e
        log(f"pe110k fp64 {out['pe110k_fp64_value']:.4e}, mixed {out['pe110k_mixed_value']:.4e} atom-steps/s")
        small = S.polyethylene(8, 12, 8)
        small.vel = S.maxwell_velocities(small, 300.0, stream=1)
        drift = {}
        for prec in ("fp64", "mixed"):
            slope, dmax = nve_drift(small, prec, 2000, 0.2e-15)
            drift[prec] = {"slope_eV_per_atom_per_ps": slope, "max_abs_dE_eV_per_atom": dmax}
        out["nve_drift"] = {"system": f"PE 8x12x8 ({small.n} atoms), 300 K, 2000 steps of 0.2 fs", **drift}
        log(f"nve drift {drift}")
        gr = S.graphite(64, 64, 32)  # 1,048,576 atoms, AB graphite, 3-D periodic
        gr.vel = S.maxwell_velocities(gr, 300.0, stream=5)
        e = Engine.from_system(gr, potentials=("REBO", "AIREBO", "ETors"), precision="mixed")
        ms_gr, _ = timed(e, a.steps, a.warmup, 0.5e-15)
        out["graphite1m_mixed_value"] = gr.n * a.steps / (ms_gr / 1e3)
        del e
        log(f"graphite 1M mixed {out['graphite1m_mixed_value']:.4e} atom-steps/s")
        out["metal_ackland"] = metal_block(a.steps, a.warmup)
        log(f"metal block (Ackland): {out['metal_ackland']['value']:.4e} atom-steps/s, "
            f"{out['metal_ackland']['kernels_ms_per_step']}")
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
