"""Impact sequences as `mdtrajsim -i` runs them, batched on one GPU, with
mdtrajsim's run control.

mdtrajsim's experiment mode (applications/mdtrajsim/main.cpp:280-348) loads a
base state, reads the projectile's lateral positions from `ionpos.bin` (raw
float64 pairs x, y in cm, one pair per impact, possibly compressed as yaatk's
stream writes it), and for every impact from the first index on puts the
last atom -- the projectile -- at that (x, y) and, in the directory `%08d` of
the impact, writes `ionpos_used.bin`/`.txt` and the initial state, then runs
the trajectory (runTraj, main.cpp:115-191) until simTime reaches
simTimeFinal.  A directory that already holds `completed.ok`,
`completed.error` or `in.mde.after_crash` is skipped (main.cpp:95-112).

Run control, as the reference does it:
* state ids are SimLoopSaver's `md` + the iteration in ten digits
  (SimLoopSaver.cxx:628-634); every write of a state or of snapshots.conf
  holds yaatk's `flag.write-is-in-progress` file while it runs (DataState,
  yaatk.hpp:435-470), and runTraj refuses a directory that still holds one
  (main.cpp:119-120);
* a trajectory starts from the latest `md%010d` state in its directory
  (main.cpp:147-152) -- a run that was cut short resumes from its last flush,
  with its accelerations, clock, iteration and snapshots.conf;
* every `iterationFlushStateInterval` = 1000 iterations (main.cpp:155) the
  state and snapshots.conf are written (SimLoop.cxx:263-268, main.cpp:84-87);
* at the end the final state and snapshots.conf; `completed.ok` when simTime
  reached simTimeFinal, then the intermediate states are removed
  (SimLoopSaver::removeIterations, SimLoopSaver.cxx:761-795: runTraj keeps
  the first and the last, the impact loop then only the last);
* a device error of one trajectory (the reference's REQUIREs, or a capacity
  of the engine) ends that trajectory alone as SimLoop::execute's catch does
  (SimLoop.cxx:176-197): `completed.error`, the state at that iteration,
  `in.mde.after_crash` (SimLoop::saveToMDE) and `mde_state.after_crash`
  (SimLoop::saveToStream); the other trajectories of the batch run on.

The trajectories of a group run as one batch in one engine context, each on
its own clock (the engine stops every system at its own simTimeFinal).
"""
from __future__ import annotations

import os
import re
import struct
from dataclasses import dataclass, field

import numpy as np

from . import state as ST

CHAIN = ("FBZL", "FLJ", "REBO", "AIREBO", "ETors", "Ackland")  # applications/common.h:46-68
FINISHED = ("completed.ok", "completed.error", "in.mde.after_crash")
FLAG = "flag.write-is-in-progress"  # yaatk::DataState::flagFilename (yaatk.cxx:639)
FLUSH_INTERVAL = 1000  # mdloop.iterationFlushStateInterval (main.cpp:155)
PS = 1e-12
AO = 1e-8
E_CGS = 4.8e-10          # mdtk/consts.cxx:29
AMU = 1.6605655e-24      # mdtk/consts.cxx:28
# Atom::setAttributesByElementID (Atom.hpp:98-113): Z (in e) and M (in amu)
ELEMENTS = {1: (1.0, 1.0), 12: (6.0, 12.0), 64: (29.0, 63.546), 108: (47.0, 107.868), 197: (79.0, 196.967),
            40: (18.0, 39.948), 131: (54.0, 131.293)}
TAG_BITS = {"target": 1 << 1, "projectile": 1 << 2, "substrate": 1 << 3, "monomer": 1 << 4, "cluster": 1 << 5,
            "fullerene": 1 << 6}  # ATOMTAG_* (Atom.hpp:84-90)


class DirtyStateError(RuntimeError):
    """A directory still holds the flag of an interrupted write."""


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


def is_clean(d: str) -> bool:
    """yaatk::DataState::isClean: no write was cut short in this directory."""
    return not os.path.exists(os.path.join(d, FLAG))


class _DataState:
    """yaatk::DataState: the flag file exists while a write runs; a flag left
    by an earlier, interrupted write is removed at the end of this one."""

    def __init__(self, d: str):
        self.path = os.path.join(d, FLAG)

    def __enter__(self):
        if not os.path.exists(self.path):
            open(self.path, "w").close()
        return self

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            os.remove(self.path)
        return False


def state_id(iteration: int) -> str:
    return f"md{int(iteration):010d}"  # SimLoopSaver::generateId


def list_iterations(d: str) -> list[int]:
    """SimLoopSaver::listIterations: the iterations of the `md<digits>` files."""
    its = set()
    for f in os.listdir(d):
        m = re.match(r"md(\d+)", f)
        if m:
            its.add(int(m.group(1)))
    return sorted(its)


def remove_iterations(d: str, keep_first: bool = True, keep_last: bool = True):
    """SimLoopSaver::removeIterations(keepFirst, keepLast)."""
    its = list_iterations(d)
    if its and keep_last:
        its = its[:-1]
    if its and keep_first:
        its = its[1:]
    ids = [state_id(i) for i in its]
    for f in os.listdir(d):
        if any(f.startswith(i) for i in ids):
            os.remove(os.path.join(d, f))


def write_state_guarded(d: str, iteration: int, st: ST.State, compress: str | None):
    """SimLoopSaver::write under yaatk::DataState."""
    with _DataState(d):
        ST.write_state(os.path.join(d, state_id(iteration)), st, compress)


def impact_state(base: ST.State, x: float, y: float) -> ST.State:
    """The base state with its last atom (the projectile) at (x, y) (main.cpp:306-308)."""
    pos = base.pos.copy()
    pos[-1, 0] = x
    pos[-1, 1] = y
    return ST.State(base.elem, pos, base.vel, base.acc, base.pbc, base.pbc_count, base.pbc_enabled, base.indices,
                    base.tags, base.bath, dict(base.time), base.check)


def _cpp_double(v: float, precision: int = 6) -> str:
    """A double as `std::ostream << v` prints it at the given precision."""
    return f"{v:.{precision}g}"


def _write_used(d: str, x: float, y: float):
    with open(os.path.join(d, "ionpos_used.bin"), "wb") as f:
        f.write(struct.pack("<2d", x, y))
    with open(os.path.join(d, "ionpos_used.txt"), "w") as f:
        f.write(f"{_cpp_double(x)} {_cpp_double(y)}\n")


# ----------------------------------------------------------- crash dumps
def _atoms_text(st: ST.State, an_no_tb, prec: int) -> list[str]:
    """Each atom as `operator<<(ostream&, const Atom&)` prints it
    (Atom.cxx:223-240), followed by YAATK_FSTREAM_WRITE's newline."""
    f = lambda v: _cpp_double(float(v), prec)  # noqa: E731
    v3 = lambda a: " ".join(f(q) for q in a) + " "  # noqa: E731 -- Vector3D (Vector3D.hpp:343-350)
    en = np.asarray(st.pbc_enabled, bool)
    out = []
    for i in range(st.n):
        e = int(st.elem[i])
        z, m = ELEMENTS.get(e, (0.0, 0.0))
        tags = sum(b for t, b in TAG_BITS.items() if st.tags.get(t, np.zeros(st.n, bool))[i])
        pbc = np.where(en[i], st.pbc, ST.NO_PBC)
        out += [str(e), f(z * E_CGS), f(m * AMU), v3(st.pos[i]), " ".join(str(int(q)) for q in st.pbc_count[i]) + " ",
                v3(st.vel[i]), v3(st.acc[i]), v3(an_no_tb[i]),
                str(int(bool(st.tags.get("thermal_bath_applicable", np.ones(st.n, bool))[i]))),
                str(int(st.indices[i])), str(int(bool(st.tags.get("fixed", np.zeros(st.n, bool))[i]))), v3(pbc),
                str(tags), ""]
    return out


def _legacy_bath_text(prec: int) -> list[str]:
    """SimLoop::LegacyThermalBathStruct at its defaults (SimLoop.cxx:1325-1338)."""
    return [_cpp_double(v, prec) for v in (1e6 * AO, 0.0, -5.0 * AO, 0.0, 1.0e13)]


def _traj_interval(sim_time: float) -> float:
    """CustomSimLoop::doBeforeIteration's simTimeSaveTrajInterval (main.cpp:58-61)."""
    return 1.0 * PS if sim_time < 2.0 * PS else 5.0 * PS


def write_after_crash(d: str, st: ST.State, iteration: int):
    """in.mde.after_crash (SimLoop::saveToMDE, SimLoop.cxx:1175-1192) and
    mde_state.after_crash (SimLoop::saveToStream, :603-629, precision 10)."""
    t = st.time
    an_nt = st.acc
    lines = [str(st.n)] + _atoms_text(st, an_nt, 6)
    lines += [_cpp_double(v) for v in (t["sim_time"], _traj_interval(t["sim_time"]), t["sim_time_final"])]
    lines += [" ".join(_cpp_double(q) for q in st.pbc) + " "] + _legacy_bath_text(6)
    with open(os.path.join(d, "in.mde.after_crash"), "w") as f:
        f.write("\n".join(lines) + "\n")
    p = 10
    c = st.check or dict(check_force=False, net_force=(0.0, 0.0, 0.0), check_energy=False, initial_energy=0.0,
                         current_energy=0.0, energy_from_bath=0.0, temperature=0.0)
    lines = [str(st.n)] + _atoms_text(st, an_nt, p) + [" ".join(_cpp_double(q, p) for q in st.pbc) + " "]
    lines += [str(int(bool(c["check_force"]))), " ".join(_cpp_double(q, p) for q in c["net_force"]) + " ",
              str(int(bool(c["check_energy"])))]
    lines += [_cpp_double(c[k], p) for k in ("initial_energy", "current_energy", "energy_from_bath", "temperature")]
    lines += [_cpp_double(v, p) for v in (t["sim_time"], _traj_interval(t["sim_time"]), t["sim_time_final"], AO,
                                          t["dt"])]
    lines += [str(int(iteration)), str(FLUSH_INTERVAL)] + _legacy_bath_text(p) + [_cpp_double(0.0, p)]
    with open(os.path.join(d, "mde_state.after_crash"), "w") as f:
        f.write("\n".join(lines) + "\n")


# ------------------------------------------------------------ the runs
@dataclass
class _Traj:
    dir: str
    state: ST.State            # the state the run starts from
    iteration: int             # its SimLoop::iteration
    resumed: bool              # loaded from a flush (accelerations kept)
    shots: list = field(default_factory=list)  # snapshots already in snapshots.conf
    sel: np.ndarray | None = None
    label: int = 0


def _start(d: str, label: int, fallback: ST.State | None = None) -> _Traj:
    """runTraj's start (main.cpp:147-152): the latest md%010d state in d."""
    its = list_iterations(d)
    if its:
        it = its[-1]
        st = ST.read_state(os.path.join(d, state_id(it)))
    else:
        it, st = 0, fallback
    tr = _Traj(d, st, it, resumed=it > 0, label=label)
    conf = os.path.join(d, "snapshots.conf")
    if os.path.exists(conf):  # SnapshotList::loadstate (main.cpp:154)
        from . import trajectory as TR
        tr.sel, tr.shots = TR.read_snapshots_conf(conf)
    return tr


def _box(st: ST.State) -> np.ndarray:
    """The system's box: the state's box on every axis some atom has enabled."""
    en = np.asarray(st.pbc_enabled, bool).reshape(-1, 3)
    return np.where(en.any(axis=0) if st.n else np.zeros(3, bool), st.pbc, ST.NO_PBC)


def _atom_pbc(st: ST.State) -> np.ndarray:
    """Each atom's own PBC vector (Atom::PBC): the box on its enabled axes."""
    en = np.asarray(st.pbc_enabled, bool).reshape(-1, 3)
    return np.where(en, st.pbc[None, :], ST.NO_PBC)


def _snap_tag(st: ST.State) -> np.ndarray:
    tag = st.tags.get("projectile", np.zeros(st.n, bool)) | st.tags.get("cluster", np.zeros(st.n, bool))
    if not tag.any():
        tag = np.eye(1, st.n, st.n - 1, dtype=bool)[0]
    return tag


def run_trajectories(trajs: list[_Traj], potentials=CHAIN, precision: str = "fp64", chunk: int = 1000,
                     compress: str | None = "xz", etors_compat: bool = True, max_steps: int = 10_000_000,
                     trajectory: bool = True, flush_interval: int = FLUSH_INTERVAL) -> list[dict]:
    """SimLoop::execute for a batch of trajectories, each in its directory
    with the run control of runTraj (flushes, final state, snapshots.conf,
    XVA frames, completed.ok / the crash files).  Stops with RuntimeError
    after max_steps engine steps (the flushed states stay, as after a kill).
    Returns per trajectory {dir, iterations, sim_time, epot, ekin, error}."""
    from .engine import Engine, FRAME_DTYPE, RECORD_DTYPE
    from . import trajectory as TR
    m = len(trajs)
    sts = [t.state for t in trajs]
    eng = Engine(potentials=potentials, precision=precision, etors_compat=etors_compat)
    eng.set_systems([(s.elem, s.pos, s.vel, _box(s), s.tags.get("fixed", np.zeros(s.n, bool)).astype(np.uint8))
                     for s in sts])
    off = eng.offsets
    pos = np.concatenate([s.pos for s in sts])
    cnt = np.concatenate([s.pbc_count for s in sts])
    if any(s.n and not (np.asarray(s.pbc_enabled, bool) == np.asarray(s.pbc_enabled, bool)[0]).all() for s in sts):
        eng.set_atom_pbc(np.concatenate([_atom_pbc(s) for s in sts]), pos, cnt)  # per-atom PBC in some state
    else:
        eng.set_positions(pos, pbc_count=cnt)
    b = sts[0].bath
    if b and b["geometry"] != 0:
        eng.set_thermal_bath(geometry=b["geometry"], To=b["To"], gamma=b["gamma"], z_min=b["z_min"],
                             d_boundary=b["d_boundary"], z_min_free=b["z_min_free"], center=b["center"],
                             radius=b["radius"])
        eng.set_bath_atoms(np.concatenate([s.tags.get("thermal_bath_applicable", np.ones(s.n, bool)) for s in sts])
                           .astype(np.uint8))
    for s, t in enumerate(trajs):
        eng.set_time(t.state.time["sim_time"], t.state.time["sim_time_final"], t.state.time["dt"], system=s)
        eng.set_iteration(t.iteration, system=s)
    keep = np.array([t.resumed for t in trajs], np.uint8)
    if keep.any():  # resumed trajectories go on from their saved accelerations (SimLoop.cxx:305-321)
        eng.set_accelerations(np.concatenate([s.acc for s in sts]), keep)
    if trajectory:
        eng.set_trajectory_output(np.concatenate([_snap_tag(s) for s in sts]).astype(np.uint8), xva=True)
    shots = [list(t.shots) for t in trajs]
    sels = [t.sel for t in trajs]
    last_e = np.zeros((m, 2))
    last_dt = np.array([t.state.time.get("dt_prev", t.state.time["dt"]) for t in trajs])
    error = np.zeros(m, np.uint32)
    done = np.zeros(m, bool)  # crashed: no further output

    def snapshot_state(s, x, v, a, c, clock):
        lo, hi = off[s], off[s + 1]
        st = sts[s]
        return ST.State(st.elem, x[lo:hi].copy(), v[lo:hi].copy(), a[lo:hi].copy(), st.pbc,
                        c[lo:hi].astype(np.int32), st.pbc_enabled, st.indices, st.tags, st.bath,
                        dict(st.time, sim_time=float(clock[s, 0]), dt=float(clock[s, 2]),
                             dt_prev=float(last_dt[s])), st.check)

    def write_snapshots(s):
        sel = sels[s] if sels[s] is not None else np.zeros(0, np.uint32)
        with _DataState(trajs[s].dir):
            TR.write_snapshots_conf(os.path.join(trajs[s].dir, "snapshots.conf"), sel, shots[s])

    steps = 0
    while True:
        clock = eng.time()
        its = eng.iterations()
        live = (clock[:, 0] < clock[:, 1]) & ~done
        if not live.any():
            break
        if steps >= max_steps:
            raise RuntimeError(f"{trajs[0].dir}..: {max_steps} steps without reaching simTimeFinal")
        # the next flush point of any live trajectory ends the call
        k = int(min(chunk, max_steps - steps, *(flush_interval - its[live] % flush_interval)))
        r = eng.run(k, allow_stopped=True)
        steps += k
        dt = r["dt"].reshape(k, m)
        ran = dt > 0
        ep, ek = r["epot"].reshape(k, m), r["ekin"].reshape(k, m)
        for s in np.nonzero(ran.any(axis=0))[0]:
            q = np.nonzero(ran[:, s])[0][-1]
            last_e[s] = ep[q, s], ek[q, s]
            last_dt[s] = dt[q, s]
        if trajectory:
            f, rc = eng.take_trajectory()
            for s in range(m):
                sel, new = TR.snapshot_list(f, rc, s)
                if len(new) and sels[s] is None:
                    sels[s] = sel
                shots[s] += new
                for fr in TR.xva_frames(f, s):  # SimLoop::writetrajXVA at the iteration's start (SimLoop.cxx:271-277)
                    with open(os.path.join(trajs[s].dir, TR.xva_name(fr)), "w") as fo:
                        fo.write(TR.xva_text(fr, rc, sts[s].check))
        clock = eng.time()
        its = eng.iterations()
        stopped = r.get("stopped")
        need = [s for s in range(m) if not done[s] and ran[:, s].any() and (
            (stopped is not None and stopped[s]) or
            (its[s] % flush_interval == 0 and clock[s, 0] < clock[s, 1]))]
        if need:
            x, v, a, c = eng.state()
        for s in need:
            st = snapshot_state(s, x, v, a, c, clock)
            if stopped is not None and stopped[s]:  # SimLoop::execute's catch (SimLoop.cxx:176-197)
                error[s] = stopped[s]
                done[s] = True
                d = trajs[s].dir
                open(os.path.join(d, "completed.error"), "w").close()
                write_state_guarded(d, its[s], st, compress)
                write_after_crash(d, st, its[s])
            else:  # the periodic flush (SimLoop.cxx:263-268, main.cpp:84-87)
                if trajectory:
                    write_snapshots(s)
                write_state_guarded(trajs[s].dir, its[s], st, compress)
    clock = eng.time()
    its = eng.iterations()
    x, v, a, c = eng.state()
    report = []
    for s, t in enumerate(trajs):
        st = snapshot_state(s, x, v, a, c, clock)
        write_state_guarded(t.dir, its[s], st, compress)  # runTraj: mds.write(); snapshotList.writestate()
        if trajectory:
            write_snapshots(s)
        if not error[s] and clock[s, 0] >= clock[s, 1]:
            open(os.path.join(t.dir, "completed.ok"), "w").close()
            remove_iterations(t.dir)  # keep the first and the last
        report.append(dict(impact=t.label, dir=t.dir, iterations=int(its[s]), sim_time=float(clock[s, 0]),
                           epot=float(last_e[s, 0]), ekin=float(last_e[s, 1]), error=int(error[s])))
    del eng
    return report


def run_impacts(base: str, ionpos: str = "ionpos.bin", out: str = ".", start: int = 0, stop: int | None = None,
                potentials=CHAIN, precision: str = "fp64", batch: int = 256, chunk: int = 1000,
                compress: str | None = "xz", etors_compat: bool = True, max_steps: int = 10_000_000,
                trajectory: bool = True, flush_interval: int = FLUSH_INTERVAL) -> list[dict]:
    """Runs impacts [start, stop) of the ionpos file from the base state
    (prefix `base` of its SimLoopSaver files) in batches of `batch`, at most
    `chunk` steps per engine call, with mdtrajsim's run control in each
    impact's directory under `out`: the initial state, resumption from the
    latest flushed state, a flush every `flush_interval` iterations, the
    final state, snapshots.conf of the atoms tagged projectile or cluster
    (the last atom when none is tagged), the XVA frames (main.cpp:155-167,
    SnapshotList.hpp:85-91, SimLoop.cxx:271-277), completed.ok, and the crash
    files of a trajectory that fails.  Returns per impact {impact, dir,
    iterations, sim_time, epot, ekin, error}."""
    st = ST.read_state(base)
    xy = read_ionpos(ionpos)
    stop = len(xy) if stop is None else min(stop, len(xy))
    todo = [k for k in range(start, stop) if not is_finished(impact_dir(out, k))]
    it0 = 0  # SimLoopSaver::load of the base leaves iteration 0 (main.cpp:286-287)
    report = []
    for g0 in range(0, len(todo), batch):
        trajs = []
        for k in todo[g0:g0 + batch]:
            d = impact_dir(out, k)
            os.makedirs(d, exist_ok=True)
            _write_used(d, xy[k, 0], xy[k, 1])  # main.cpp:315-327
            s0 = impact_state(st, xy[k, 0], xy[k, 1])
            # mds.write() (main.cpp:337): its DataState also clears a flag left
            # by an interrupted write, so runTraj's isClean test (main.cpp:119)
            # passes in this mode
            write_state_guarded(d, it0, s0, compress)
            trajs.append(_start(d, k, s0))
        if trajs:
            rep = run_trajectories(trajs, potentials, precision, chunk, compress, etors_compat, max_steps,
                                   trajectory, flush_interval)
            for r in rep:
                remove_iterations(r["dir"], keep_first=False, keep_last=True)  # main.cpp:341
            report += rep
    return report


def run_common(input_id: str = "base", out: str = ".", potentials=CHAIN, precision: str = "fp64",
               chunk: int = 1000, compress: str | None = "xz", etors_compat: bool = True,
               max_steps: int = 10_000_000, flush_interval: int = FLUSH_INTERVAL) -> dict | None:
    """mdtrajsim's common usage (runTraj(inputFilesId), main.cpp:115-191):
    nothing when `out` holds a finished run; DirtyStateError when a write
    there was cut short; otherwise the state `input_id` is loaded, written
    as md0000000000 and run to its simTimeFinal in `out` with the periodic
    flush, the final state, snapshots.conf, the XVA frames and completed.ok
    (the first and the last state stay)."""
    if is_finished(out):
        return None
    os.makedirs(out, exist_ok=True)
    if not is_clean(out):
        raise DirtyStateError(f"{out}: {FLAG} is present (an interrupted write); the state there is not trusted")
    st = ST.read_state(input_id)
    write_state_guarded(out, 0, st, compress)
    tr = _Traj(out, st, 0, resumed=False)
    return run_trajectories([tr], potentials, precision, chunk, compress, etors_compat, max_steps, True,
                            flush_interval)[0]


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
        try:
            r = run_common(a.file, ".", precision=a.precision)
        except DirtyStateError as e:
            print(e, flush=True)
            return 1
        rep = [] if r is None else [r]
    rc = 0
    for r in rep:
        print(f"{r['dir']}: {r['iterations']} iterations, simTime {r['sim_time']:.6e} s"
              + (f", device error bits {r['error']}" if r["error"] else ""), flush=True)
        rc |= 1 if r["error"] else 0
    return rc


if __name__ == "__main__":
    raise SystemExit(main())
