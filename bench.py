"""Benchmark: AIREBO NVE steps of a polyethylene crystal on one MI355X.

Workload (one "step" = one velocity-Verlet iteration: drift, skin test and
neighbour-list rebuild when due, REBO + AIREBO-LJ + torsion force call, kick,
PBC wrap): orthorhombic polyethylene, physical two-chain Bunn cell
(a=7.42, b=4.95, c=2.55 A, 12 atoms), 24x24x16 cells = 110,592 atoms,
3-D periodic, 300 K Maxwell-Boltzmann velocities (Philox stream 0, seed
697860), dt = 0.2 fs, reference parameter set (LJ switched off 5.0-5.5 A).
Synthetic input; positions and velocities are resident in HBM before the
timed region.

Prints ONE JSON line (rank 0).  `value` = atoms x steps / wall seconds of the
timed region (hipEvents on the engine's stream, bracketed by device
synchronisation).  `roofline` prices the dominant kernel against MI355X's FP64
peak; `cpu_baseline` times the oracle (the CPU restatement of the reference's
algorithm, single thread) on a bounded sample of the same workload.
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
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# Algorithmic FP64 work per evaluation (forward + adjoint), tallied from the
# kernel source with FMA=2, add/mul=1, div/sqrt=8, exp/sin/cos=20, pow=40
# flops; per-neighbour terms use the measured neighbour counts of the workload.
FLOP_BOND_ORDER_BASE = 900.0       # pair set-up, coordination, pi_rc tricubic, p_ij pow, Nconj
FLOP_PER_SIGMA_NEIGHBOUR = 260.0   # cos theta, G quintic, cutoff/ExpTerm, adjoint scatter (per k, both passes)
FLOP_LJ_PAIR = 120.0               # cutoff, sigma/r powers, C_ij merge, gradient


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=str, default="24,24,16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


def cpu_baseline(budget_s: float):
    """Oracle (single-threaded CPU restatement) on a 4x6x10-cell PE sample."""
    import numpy as np
    from mdtk_amd import systems as S
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO, POT_ETORS, build
    build()
    s = S.polyethylene(4, 6, 10)
    s.vel = S.maxwell_velocities(s, 300.0, stream=0)
    o = Oracle.from_system(s, pots=POT_REBO | POT_AIREBO | POT_ETORS)
    o.run(1, fixed_dt=0.2e-15)  # prepare + first step (NL build, initial force) untimed
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.run(1, fixed_dt=0.2e-15)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": s.n * steps / dt, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"polyethylene 4x6x10 Bunn cells ({s.n} atoms), REBO+AIREBO-LJ+ETors FP64, "
                      f"dt=0.2 fs, {steps} steps in {dt:.1f} s, single thread"}


KERNELS = ("k_bstar", "k_rebo_bonds", "k_lj_wave", "k_etors", "k_coord", "k_bond_enum", "k_nl_build_long",
           "k_nl_build_short", "k_kick", "k_drift", "k_disp_top2", "k_disp_decide", "k_zero")


def measure(sysm, precision, steps, warmup, dt):
    """Times `steps` Verlet iterations (one hipGraph launch each) between two
    HIP events on the engine's stream, then repeats them with per-kernel event
    timing for the roofline."""
    import torch
    from mdtk_amd import Engine
    eng = Engine.from_system(sysm, potentials=("REBO", "AIREBO", "ETors"), precision=precision)
    stream = torch.cuda.Stream()
    eng.set_stream(stream.cuda_stream)
    eng.run(max(warmup, 1), fixed_dt=dt)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    res = eng.run(steps, fixed_dt=dt)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    eng.profile(True)
    eng.run(steps, fixed_dt=dt)
    kt = {k: eng.kernel_time(k) for k in KERNELS}
    eng.profile(False)
    cnt = eng.counters()
    return eng, ms, kt, cnt, res


def roofline(kt, cnt):
    dom = max(kt, key=lambda k: kt[k][0])
    dom_ms, dom_n = kt[dom]
    avg_s = dom_ms / max(dom_n, 1) / 1e3
    n_sigma = 2.0 * 3.3  # sigma-pi neighbours visited per bond order (both sides), PE average
    per_bo = FLOP_BOND_ORDER_BASE + FLOP_PER_SIGMA_NEIGHBOUR * n_sigma
    if dom == "k_bstar":
        flops = cnt["bstar"] * (per_bo + FLOP_LJ_PAIR)
    elif dom == "k_rebo_bonds":
        flops = cnt["bonds"] * (per_bo + 200.0)
    else:
        flops = cnt["lj_pairs"] * FLOP_LJ_PAIR
    return dom, dom_n, avg_s, flops / avg_s / 1e12


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU per process; --gpus {a.gpus} is not supported "
              "(independent trajectories shard with no communication; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from mdtk_amd import systems as S

    nx, ny, nz = (int(v) for v in a.cells.split(","))
    sysm = S.polyethylene(nx, ny, nz)
    sysm.vel = S.maxwell_velocities(sysm, 300.0, stream=0)
    dt = 0.2e-15
    # force error of the mixed form against FP64 at the initial configuration
    from mdtk_amd import Engine
    e64 = Engine.from_system(sysm, potentials=("REBO", "AIREBO", "ETors"))
    emx = Engine.from_system(sysm, potentials=("REBO", "AIREBO", "ETors"), precision="mixed")
    E64, g64 = e64.energy_grad()
    Emx, gmx = emx.energy_grad()
    ev_per_a = S.EV / S.ANG
    f_rms = float(np.sqrt(((g64 - gmx) ** 2).sum(1).mean()) / ev_per_a)
    del e64, emx
    _, ms64, kt64, cnt64, _ = measure(sysm, "fp64", a.steps, a.warmup, dt)
    _, msmx, ktmx, cntmx, res = measure(sysm, "mixed", a.steps, a.warmup, dt)
    v64 = sysm.n * a.steps / (ms64 / 1e3)
    vmx = sysm.n * a.steps / (msmx / 1e3)
    use_mixed = f_rms <= 1e-5
    value, ms, kt, cnt = (vmx, msmx, ktmx, cntmx) if use_mixed else (v64, ms64, kt64, cnt64)
    dom, dom_n, avg_s, achieved = roofline(kt, cnt)
    peak = FP32_PEAK_TFLOPS if use_mixed else FP64_PEAK_TFLOPS
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
        "config": {"workload": f"AIREBO NVE polyethylene {nx}x{ny}x{nz} Bunn cells (2 chains/cell)",
                   "atoms": sysm.n, "dt_fs": 0.2, "potentials": "REBO+AIREBO-LJ+ETors (reference cutoffs)",
                   "precision": "mixed" if use_mixed else "fp64", "batched": False,
                   "parallelism": "single GPU"},
        "fp64_value": v64,
        "fp64_ms_per_step": ms64 / a.steps,
        "mixed_value": vmx,
        "mixed_over_fp64": vmx / v64,
        "mixed_force_rms_err_eV_per_A": f_rms,
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "avg_launch_ms": avg_s * 1e3, "launches": dom_n,
                     "note": ("vector-ALU peak of the compute type (no dense contraction on this path: "
                              "no MFMA)")},
        "kernels_ms_per_step": {k: v[0] / max(a.steps, 1) for k, v in kt.items()},
        "kernels_ms_per_step_fp64": {k: v[0] / max(a.steps, 1) for k, v in kt64.items()},
        "counters": cnt,
        "energy_check": {"etot_first_eV": float((res["epot"][0] + res["ekin"][0]) / S.EV),
                         "etot_last_eV": float((res["epot"][-1] + res["ekin"][-1]) / S.EV)},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
