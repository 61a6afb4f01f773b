"""Trajectory files from frames captured on the GPU (Engine.set_trajectory_output).

* `snapshots.conf` -- mdtrajsim's SnapshotList (mdtk/SnapshotList.hpp:38-204):
  uint32 count of selected atoms, their uint32 indices, uint32 count of
  snapshots, then per snapshot float64 simTime, uint32 count and one 36-byte
  AtomSnapshot per selected atom (int32 PBC_count[3], float32 pos[3] (cm),
  float32 v[3] (cm/s)), all little-endian.
* `mde%010d.xva` -- SimLoop::saveToStreamXVA (mdtk/SimLoop.cxx:681-748): the
  velocity scale 10^(ceil(log10(mean |v_i|)) - 2), the distance scale 1 A
  (cm), the atom count, per atom v/scale, x/A (fixed, 2 decimals) and the
  PBC counters, then the check record, simTime, dt and iteration in
  scientific notation with 10 digits, and the CPU time (0 here).
"""
from __future__ import annotations

import math
import struct

import numpy as np

from .engine import FRAME_SNAPSHOT, FRAME_XVA

ATOM_SNAPSHOT = np.dtype([("pbc_count", "<i4", 3), ("pos", "<f4", 3), ("v", "<f4", 3)])  # 36 bytes


def snapshot_list(frames, records, system: int = 0):
    """(selected atom indices, [(simTime, AtomSnapshot array)]) of one system."""
    shots, sel = [], None
    for f in frames[(frames["kind"] == FRAME_SNAPSHOT) & (frames["system"] == system)]:
        r = records[f["first"]:f["first"] + f["count"]]
        a = np.zeros(len(r), ATOM_SNAPSHOT)
        a["pbc_count"] = r["pbc_count"]
        a["pos"] = r["x"].astype(np.float32)  # AtomSnapshot(const Atom&): double -> float
        a["v"] = r["v"].astype(np.float32)
        if sel is None:
            sel = r["atom"].astype(np.uint32)
        shots.append((float(f["sim_time"]), a))
    return (np.zeros(0, np.uint32) if sel is None else sel), shots


def write_snapshots_conf(path: str, selected, shots):
    """SnapshotList::writestate (SnapshotList.hpp:133-165)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(selected)))
        f.write(np.asarray(selected, "<u4").tobytes())
        f.write(struct.pack("<I", len(shots)))
        for t, a in shots:
            f.write(struct.pack("<dI", t, len(a)))
            f.write(np.ascontiguousarray(a, ATOM_SNAPSHOT).tobytes())


def read_snapshots_conf(path: str):
    """SnapshotList::loadstate: (selected, [(simTime, AtomSnapshot array)])."""
    with open(path, "rb") as f:
        d = f.read()
    (n,), o = struct.unpack_from("<I", d, 0), 4
    sel = np.frombuffer(d, "<u4", n, o).copy()
    o += 4 * n
    (k,), o = struct.unpack_from("<I", d, o), o + 4
    shots = []
    for _ in range(k):
        t, m = struct.unpack_from("<dI", d, o)
        o += 12
        shots.append((t, np.frombuffer(d, ATOM_SNAPSHOT, m, o).copy()))
        o += m * ATOM_SNAPSHOT.itemsize
    if o != len(d):
        raise ValueError(f"{path}: {len(d) - o} trailing bytes")
    return sel, shots


def xva_text(frame, records, check=None) -> str:
    """One XVA frame as SimLoop::saveToStreamXVA prints it."""
    r = records[frame["first"]:frame["first"] + frame["count"]]
    v, x, pc = r["v"], r["x"], r["pbc_count"]
    scale = float(np.abs(v).sum() / (3.0 * len(r))) if len(r) else 0.0
    if abs(scale) < 1e-30:
        scale = 1e-30
    scale = 10.0 ** (math.ceil(math.log10(abs(scale))) - 2)
    ao = 1e-8
    out = [f"{scale:g}", f"{ao:g}", str(len(r))]
    for i in range(len(r)):
        out.append(" ".join(f"{q / scale:.2f}" for q in v[i]))
        out.append(" ".join(f"{q / ao:.2f}" for q in x[i]))
        out.append(" ".join(str(int(q)) for q in pc[i]))
    c = check or dict(check_force=True, net_force=(0.0, 0.0, 0.0), check_energy=True, initial_energy=0.0,
                      current_energy=0.0, energy_from_bath=0.0, temperature=0.0)
    e = lambda q: f"{q:.10e}"  # noqa: E731
    out.append(str(int(bool(c["check_force"]))))
    out.append(" ".join(e(q) for q in c["net_force"]) + " ")
    out.append(str(int(bool(c["check_energy"]))))
    for k in ("initial_energy", "current_energy", "energy_from_bath", "temperature"):
        out.append(e(c[k]))
    out += [e(frame["sim_time"]), e(frame["dt"]), str(int(frame["iteration"])), e(0.0)]
    return "\n".join(out) + "\n"


def xva_name(frame) -> str:
    return f"mde{int(frame['iteration']):010d}.xva"  # SimLoop::writetrajXVA (SimLoop.cxx:821-832)


def xva_frames(frames, system: int = 0):
    return frames[(frames["kind"] == FRAME_XVA) & (frames["system"] == system)]
