"""Impact sequences as `mdtrajsim -i` runs them, batched on one GPU.

mdtrajsim's experiment mode (applications/mdtrajsim/main.cpp:280-348) loads a
base state, reads the projectile's lateral positions from `ionpos.bin` (raw
float64 pairs x, y in cm, one pair per impact, possibly compressed as yaatk's
stream writes it), and for every impact from the first index on puts the
last atom -- the projectile -- at that (x, y), runs the trajectory in the
directory `%08d` of the impact until simTime reaches simTimeFinal, and
leaves there `ionpos_used.bin`/`.txt`, the final state and `completed.ok`
(a directory that already holds `completed.ok`, `completed.error` or
`in.mde.after_crash` is skipped, main.cpp:95-112).

Here the impacts of a group run as one batch in one engine context, each on
its own clock (the engine stops every system at its own simTimeFinal), and
each impact's directory receives what the reference leaves there.  The
final state's id is SimLoopSaver's `md` + the iteration number in ten digits
(SimLoopSaver.cxx:628-634).
"""
from __future__ import annotations

import os
import struct

import numpy as np

from . import state as ST

CHAIN = ("FBZL", "FLJ", "REBO", "AIREBO", "ETors", "Ackland")  # applications/common.h:46-68
FINISHED = ("completed.ok", "completed.error", "in.mde.after_crash")


def read_ionpos(path: str) -> np.ndarray:
    """(k, 2) projectile positions (cm) from an ionpos.bin (plain, .xz or .gz)."""
    d = ST._read(path)
    if d is None:
        raise FileNotFoundError(path)
    if len(d) % 16:  # main.cpp:295-297
        raise ValueError(f"{path}: {len(d)} bytes is not a whole number of (x, y) float64 pairs")
    return np.frombuffer(d, "<f8").reshape(-1, 2).copy()


def write_ionpos(path: str, xy, compress: str | None = None):
    ST._write(path, np.ascontiguousarray(xy, "<f8").tobytes(), compress)


def impact_dir(out: str, k: int) -> str:
    return os.path.join(out, f"{k:08d}")


def is_finished(d: str) -> bool:
    return any(os.path.exists(os.path.join(d, f)) for f in FINISHED)


def impact_state(base: ST.State, x: float, y: float) -> ST.State:
    """The base state with its last atom (the projectile) at (x, y) (main.cpp:306-308)."""
    pos = base.pos.copy()
    pos[-1, 0] = x
    pos[-1, 1] = y
    return ST.State(base.elem, pos, base.vel, base.acc, base.pbc, base.pbc_count, base.pbc_enabled, base.indices,
                    base.tags, base.bath, dict(base.time), base.check)


def _cpp_double(v: float) -> str:
    """A double as `std::ostream << v` prints it (6 significant digits)."""
    return f"{v:g}"


def _write_used(d: str, x: float, y: float):
    with open(os.path.join(d, "ionpos_used.bin"), "wb") as f:
        f.write(struct.pack("<2d", x, y))
    with open(os.path.join(d, "ionpos_used.txt"), "w") as f:
        f.write(f"{_cpp_double(x)} {_cpp_double(y)}\n")


def run_impacts(base: str, ionpos: str = "ionpos.bin", out: str = ".", start: int = 0, stop: int | None = None,
                potentials=CHAIN, precision: str = "fp64", batch: int = 256, chunk: int = 1000,
                compress: str | None = "xz", etors_compat: bool = True, max_steps: int = 10_000_000,
                trajectory: bool = True) -> list[dict]:
    """Runs impacts [start, stop) of the ionpos file from the base state
    (prefix `base` of its SimLoopSaver files) in batches of `batch`, `chunk`
    steps per engine call, and writes each impact's directory under `out`;
    with `trajectory`, also the SnapshotList of the atoms tagged projectile
    or cluster (the last atom when none is tagged) as snapshots.conf and the
    XVA frames as mde%010d.xva, as mdtrajsim's run leaves them
    (main.cpp:155-167, SnapshotList.hpp:85-91, SimLoop.cxx:271-277).
    Returns per impact {impact, dir, iterations, sim_time, epot, ekin}."""
    from .engine import Engine
    st = ST.read_state(base)
    xy = read_ionpos(ionpos)
    stop = len(xy) if stop is None else min(stop, len(xy))
    todo = [k for k in range(start, stop) if not is_finished(impact_dir(out, k))]
    en = np.asarray(st.pbc_enabled, bool)
    if st.n and not (en == en[0]).all():
        raise ValueError("atoms with different PBC axes in one state: the engine takes one box per system")
    pbc = np.where(en[0] if st.n else np.zeros(3, bool), st.pbc, ST.NO_PBC)
    fixed = st.tags.get("fixed", np.zeros(st.n, bool)).astype(np.uint8)
    t = st.time
    report = []
    for g0 in range(0, len(todo), batch):
        group = todo[g0:g0 + batch]
        states = [impact_state(st, xy[k, 0], xy[k, 1]) for k in group]
        eng = Engine(potentials=potentials, precision=precision, etors_compat=etors_compat)
        eng.set_systems([(s.elem, s.pos, s.vel, pbc, fixed) for s in states])
        off = eng.offsets
        x0 = np.concatenate([s.pos for s in states])
        c0 = np.concatenate([s.pbc_count for s in states])
        eng.set_positions(x0, pbc_count=c0)
        if st.bath and st.bath["geometry"] != 0:
            b = st.bath
            eng.set_thermal_bath(geometry=b["geometry"], To=b["To"], gamma=b["gamma"], z_min=b["z_min"],
                                 d_boundary=b["d_boundary"], z_min_free=b["z_min_free"], center=b["center"],
                                 radius=b["radius"])
            app = st.tags.get("thermal_bath_applicable", np.ones(st.n, bool)).astype(np.uint8)
            eng.set_bath_atoms(np.tile(app, len(group)))
        eng.set_time(t["sim_time"], t["sim_time_final"], t["dt"])
        if trajectory:
            tag = st.tags.get("projectile", np.zeros(st.n, bool)) | st.tags.get("cluster", np.zeros(st.n, bool))
            if not tag.any():
                tag = np.eye(1, st.n, st.n - 1, dtype=bool)[0]
            eng.set_trajectory_output(np.tile(tag.astype(np.uint8), len(group)), xva=True)
            frames, recs = [], []
        its = np.zeros(len(group), np.int64)
        last_e = np.zeros((len(group), 2))
        steps = 0
        while True:
            clock = eng.time()
            if (clock[:, 0] >= clock[:, 1]).all():
                break
            if steps >= max_steps:
                raise RuntimeError(f"impacts {group[0]}..{group[-1]}: {max_steps} steps without reaching simTimeFinal")
            r = eng.run(chunk)
            steps += chunk
            if trajectory:
                f, rc = eng.take_trajectory()
                f["first"] += sum(len(q) for q in recs)
                frames.append(f)
                recs.append(rc)
            live = r["dt"].reshape(chunk, len(group)) > 0
            its += live.sum(axis=0)
            ep, ek = r["epot"].reshape(chunk, -1), r["ekin"].reshape(chunk, -1)
            for s in range(len(group)):
                if live[:, s].any():
                    q = np.nonzero(live[:, s])[0][-1]
                    last_e[s] = ep[q, s], ek[q, s]
        clock = eng.time()
        x, v, a, c = eng.state()
        if trajectory:
            from . import trajectory as TR
            from .engine import FRAME_DTYPE, RECORD_DTYPE
            fa = np.concatenate(frames) if frames else np.zeros(0, FRAME_DTYPE)
            ra = np.concatenate(recs) if recs else np.zeros(0, RECORD_DTYPE)
        for s, k in enumerate(group):
            d = impact_dir(out, k)
            os.makedirs(d, exist_ok=True)
            _write_used(d, xy[k, 0], xy[k, 1])
            lo, hi = off[s], off[s + 1]
            fin = ST.State(states[s].elem, x[lo:hi], v[lo:hi], a[lo:hi], st.pbc, c[lo:hi].astype(np.int32),
                           st.pbc_enabled, st.indices, st.tags, st.bath,
                           dict(t, sim_time=float(clock[s, 0]), dt=float(clock[s, 2])), st.check)
            ST.write_state(os.path.join(d, f"md{int(t.get('iteration', 0)) + int(its[s]):010d}"), fin, compress)
            if trajectory:
                sel, shots = TR.snapshot_list(fa, ra, s)
                TR.write_snapshots_conf(os.path.join(d, "snapshots.conf"), sel, shots)
                for f in TR.xva_frames(fa, s):
                    with open(os.path.join(d, TR.xva_name(f)), "w") as fo:
                        fo.write(TR.xva_text(f, ra, st.check))
            with open(os.path.join(d, "completed.ok"), "w"):
                pass
            report.append(dict(impact=k, dir=d, iterations=int(its[s]), sim_time=float(clock[s, 0]),
                               epot=float(last_e[s, 0]), ekin=float(last_e[s, 1])))
        del eng
    return report


def run_common(input_id: str = "base", out: str = ".", potentials=CHAIN, precision: str = "fp64",
               chunk: int = 1000, compress: str | None = "xz", etors_compat: bool = True) -> dict | None:
    """mdtrajsim's common usage (main.cpp:115-191): one trajectory from the
    state `input_id` to its simTimeFinal, in `out`; nothing when `out` already
    holds a finished run.  Writes the final state, snapshots.conf, the XVA
    frames and completed.ok there."""
    if is_finished(out):
        return None
    os.makedirs(out, exist_ok=True)
    xy = read_state_projectile(input_id)
    path = os.path.join(out, ".ionpos_common.bin")
    write_ionpos(path, xy[None, :])
    try:
        rep = run_impacts(input_id, path, out=os.path.join(out, ".common"), potentials=potentials,
                          precision=precision, batch=1, chunk=chunk, compress=compress, etors_compat=etors_compat)
    finally:
        os.remove(path)
    d = impact_dir(os.path.join(out, ".common"), 0)
    for f in os.listdir(d):
        if not f.startswith("ionpos_used"):
            os.replace(os.path.join(d, f), os.path.join(out, f))
    for f in os.listdir(d):
        os.remove(os.path.join(d, f))
    os.rmdir(d)
    os.rmdir(os.path.join(out, ".common"))
    r = rep[0]
    r["dir"] = out
    return r


def read_state_projectile(input_id: str) -> np.ndarray:
    """(x, y) of the state's last atom, so that a common run keeps it in place."""
    st = ST.read_state(input_id)
    return st.pos[-1, :2].copy()


def main(argv=None):
    """`python -m mdtk_amd.impacts [-c | -i N] [FILE]`: mdtrajsim's two modes
    (main.cpp:193-348) on the GPU; FILE is the state id (default base)."""
    import argparse
    ap = argparse.ArgumentParser(prog="python -m mdtk_amd.impacts")
    ap.add_argument("file", nargs="?", default="base")
    ap.add_argument("-c", "--common-usage", action="store_true")
    ap.add_argument("-i", "--start-from-impact", type=int, default=None)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--batch", type=int, default=256)
    a = ap.parse_args(argv)
    if a.start_from_impact is not None and not a.common_usage:
        if a.start_from_impact < 0:
            ap.error("Wrong starting impact number")
        rep = run_impacts(a.file, "ionpos.bin", ".", start=a.start_from_impact, precision=a.precision,
                          batch=a.batch)
    else:
        r = run_common(a.file, ".", precision=a.precision)
        rep = [] if r is None else [r]
    for r in rep:
        print(f"{r['dir']}: {r['iterations']} iterations, simTime {r['sim_time']:.6e} s", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

