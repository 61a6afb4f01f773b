"""Simulation state files in the reference's SimLoopSaver layout.

mdtrajsim keeps a simulation's state as one file per attribute, `<id>.<attr>`
(mdtk/SimLoopSaver.cxx:117-230 write, :232-625 load), each file a raw
little-endian array over the atoms (or a few scalars), optionally compressed
by yaatk's stream as `<file>.xz` or `<file>.gz` (yaatk/yaatk.cxx:104-178):

  .z                     uint8 element id per atom
  .r .v .a               3 float64 per atom (cm, cm/s, cm/s^2)
  .pbc_rect              3 float64 (the array's PBC box, cm)
  .pbc_rect.count        3 int32 per atom (Atom::PBC_count)
  .pbc_rect.enabled      3 uint8 per atom (PBC per axis enabled)
  .indices               uint32 per atom (Atom::globalIndex)
  .tag.<name>            uint8 per atom for thermal_bath_applicable, fixed,
                         target, projectile, substrate, monomer, cluster,
                         fullerene
  .thermal_bath.common   float64 To, gamma (only with a bath)
  .thermal_bath.univserse  empty (the reference's spelling)
  .thermal_bath.box      float64 zMin, dBoundary, zMinOfFreeZone
  .thermal_bath.sphere   float64 center[3], radius, zMinOfFreeZone
  .time                  float64 simTime, simTimeFinal, dt, dt_prev
  .check                 uint8 checkForce, float64 netForce[3], uint8
                         checkEnergy, float64 initialEnergy, currentEnergy,
                         energyTransferredFromBath, currentTemperature

`read_state` / `write_state` move such a state to and from a `State`;
`engine_from_state` starts an engine on it and `update_state` copies an
engine's positions, velocities, accelerations and PBC counters back.
"""
from __future__ import annotations

import gzip
import lzma
import os
import struct
from dataclasses import dataclass, field

import numpy as np

NO_PBC = 10000.0 * 1e-8  # cm (mdtk/consts.hpp:96)
TAGS = ("thermal_bath_applicable", "fixed", "target", "projectile", "substrate", "monomer", "cluster",
        "fullerene")
BATH_FILES = {1: "univserse", 2: "box", 3: "sphere"}


@dataclass
class State:
    elem: np.ndarray                 # int32 [n]
    pos: np.ndarray                  # float64 [n,3] cm
    vel: np.ndarray                  # float64 [n,3] cm/s
    acc: np.ndarray                  # float64 [n,3] cm/s^2
    pbc: np.ndarray                  # float64 [3] cm (the array's box)
    pbc_count: np.ndarray            # int32 [n,3]
    pbc_enabled: np.ndarray          # bool [n,3]
    indices: np.ndarray              # uint32 [n]
    tags: dict = field(default_factory=dict)   # name -> bool [n]
    bath: dict | None = None         # geometry (1-3), To, gamma, z_min, d_boundary, z_min_free, center, radius
    # SimLoop's defaults (SimLoop.cxx:46-50): simTime 0, simTimeFinal 4 ps, dt 1e-20 s
    time: dict = field(default_factory=lambda: dict(sim_time=0.0, sim_time_final=4e-12, dt=1e-20, dt_prev=1e-20))
    check: dict | None = None

    @property
    def n(self) -> int:
        return int(self.elem.shape[0])


def _read(path: str) -> bytes | None:
    """The file's bytes, decompressed when it exists as .xz or .gz
    (yaatk::Stream::guessZipTypeByPresence, yaatk.cxx:161-178)."""
    for ext, opener in ((".xz", lzma.open), (".gz", gzip.open), ("", open)):
        if os.path.exists(path + ext):
            with opener(path + ext, "rb") as f:
                return f.read()
    return None


def _write(path: str, data: bytes, compress: str | None):
    if compress == "xz":
        with lzma.open(path + ".xz", "wb", format=lzma.FORMAT_XZ, check=lzma.CHECK_CRC64) as f:
            f.write(data)
    elif compress == "gz":
        with gzip.open(path + ".gz", "wb") as f:
            f.write(data)
    elif compress is None:
        with open(path, "wb") as f:
            f.write(data)
    else:
        raise ValueError(f"compression {compress!r}: 'xz', 'gz' or None")


def _per_atom(data: bytes, dtype, width: int, n: int | None, name: str) -> np.ndarray:
    a = np.frombuffer(data, dtype=dtype)
    if a.size % width:
        raise ValueError(f"{name}: {len(data)} bytes is not a whole number of records")
    a = a.reshape(-1, width) if width > 1 else a
    if n is not None and a.shape[0] != n:  # SimLoopSaver::prepareForAttributeReading
        raise ValueError(f"{name}: {a.shape[0]} atoms, expected {n}")
    return a.copy()


def read_state(prefix: str) -> State:
    z = _read(prefix + ".z")
    r = _read(prefix + ".r")
    if z is None or r is None:
        raise FileNotFoundError(f"{prefix}.z and {prefix}.r are needed")
    elem = _per_atom(z, "<u1", 1, None, ".z").astype(np.int32)
    n = elem.shape[0]
    pos = _per_atom(r, "<f8", 3, n, ".r")

    def opt(ext, dtype, width, default):
        d = _read(prefix + ext)
        return default if d is None else _per_atom(d, dtype, width, n, ext)

    vel = opt(".v", "<f8", 3, np.zeros((n, 3)))
    acc = opt(".a", "<f8", 3, np.zeros((n, 3)))
    d = _read(prefix + ".pbc_rect")
    pbc = np.full(3, NO_PBC) if d is None else np.frombuffer(d, "<f8").copy()
    if pbc.shape != (3,):
        raise ValueError(".pbc_rect: 3 float64 expected")
    cnt = opt(".pbc_rect.count", "<i4", 3, np.zeros((n, 3), np.int32))
    en = opt(".pbc_rect.enabled", "<u1", 3, np.tile(pbc != NO_PBC, (n, 1)).astype(np.uint8)).astype(bool)
    idx = opt(".indices", "<u4", 1, np.arange(n, dtype=np.uint32))
    tags = {}
    for t in TAGS:
        v = _read(prefix + ".tag." + t)
        tags[t] = (np.ones(n, bool) if t == "thermal_bath_applicable" else np.zeros(n, bool)) if v is None \
            else _per_atom(v, "<u1", 1, n, ".tag." + t).astype(bool)
    bath = None
    common = _read(prefix + ".thermal_bath.common")
    if common is not None:
        if len(common) != 16:
            raise ValueError(".thermal_bath.common: 2 float64 expected")
        To, gamma = struct.unpack("<2d", common)
        bath = dict(geometry=0, To=To, gamma=gamma, z_min=1e6 * 1e-8, d_boundary=0.0, z_min_free=-5e-8,
                    center=(0.0, 0.0, 0.0), radius=1e6 * 1e-8)
        if _read(prefix + ".thermal_bath.univserse") is not None:
            bath["geometry"] = 1
        b = _read(prefix + ".thermal_bath.box")
        if b is not None:
            bath.update(geometry=2, **dict(zip(("z_min", "d_boundary", "z_min_free"), struct.unpack("<3d", b))))
        s = _read(prefix + ".thermal_bath.sphere")
        if s is not None:
            cx, cy, cz, rad, zf = struct.unpack("<5d", s)
            bath.update(geometry=3, center=(cx, cy, cz), radius=rad, z_min_free=zf)
    st = State(elem, pos, vel, acc, pbc, cnt.astype(np.int32), en, idx.astype(np.uint32), tags, bath)
    t = _read(prefix + ".time")
    if t is not None:
        st.time = dict(zip(("sim_time", "sim_time_final", "dt", "dt_prev"), struct.unpack("<4d", t)))
    c = _read(prefix + ".check")
    if c is not None:
        vals = struct.unpack("<B3dB4d", c)
        st.check = dict(check_force=bool(vals[0]), net_force=vals[1:4], check_energy=bool(vals[4]),
                        initial_energy=vals[5], current_energy=vals[6], energy_from_bath=vals[7],
                        temperature=vals[8])
    return st


def write_state(prefix: str, st: State, compress: str | None = "xz"):
    """Writes every file SimLoopSaver::write writes (SimLoopSaver.cxx:117-230)."""
    n = st.n
    w = lambda ext, data: _write(prefix + ext, data, compress)  # noqa: E731
    w(".z", np.asarray(st.elem, np.uint8).tobytes())
    w(".r", np.ascontiguousarray(st.pos, "<f8").tobytes())
    w(".v", np.ascontiguousarray(st.vel, "<f8").tobytes())
    w(".pbc_rect", np.ascontiguousarray(st.pbc, "<f8").tobytes())
    w(".pbc_rect.count", np.ascontiguousarray(st.pbc_count, "<i4").tobytes())
    w(".pbc_rect.enabled", np.ascontiguousarray(st.pbc_enabled, np.uint8).tobytes())
    w(".a", np.ascontiguousarray(st.acc, "<f8").tobytes())
    w(".indices", np.ascontiguousarray(st.indices, "<u4").tobytes())
    for t in TAGS:
        default = np.ones(n, bool) if t == "thermal_bath_applicable" else np.zeros(n, bool)
        w(".tag." + t, np.asarray(st.tags.get(t, default), np.uint8).tobytes())
    b = st.bath
    if b and b["geometry"] != 0:
        w(".thermal_bath.common", struct.pack("<2d", b["To"], b["gamma"]))
        g = b["geometry"]
        if g == 1:
            w(".thermal_bath.univserse", b"")
        elif g == 2:
            w(".thermal_bath.box", struct.pack("<3d", b["z_min"], b["d_boundary"], b["z_min_free"]))
        elif g == 3:
            w(".thermal_bath.sphere", struct.pack("<5d", *b["center"], b["radius"], b["z_min_free"]))
    t = st.time
    w(".time", struct.pack("<4d", t["sim_time"], t["sim_time_final"], t["dt"], t["dt_prev"]))
    c = st.check or dict(check_force=False, net_force=(0.0, 0.0, 0.0), check_energy=False, initial_energy=0.0,
                         current_energy=0.0, energy_from_bath=0.0, temperature=0.0)
    w(".check", struct.pack("<B3dB4d", int(c["check_force"]), *c["net_force"], int(c["check_energy"]),
                            c["initial_energy"], c["current_energy"], c["energy_from_bath"], c["temperature"]))


def engine_from_state(st: State, clock: bool = True, **engine_kw):
    """An Engine on the state's atoms: elements, positions (with their PBC
    counters), velocities, fixed atoms, the thermal bath and its per-atom
    opt-in, and (clock=True) SimLoop's clock from the .time record: simTime,
    simTimeFinal and dt, so that run() stops where the reference's loop would.
    Each atom's PBC is the box on its enabled axes and NO_PBC on the others
    (Atom::PBC as SimLoopSaver stores it); atoms whose axes differ take
    mdg_set_atom_pbc.  Accelerations start from the forces at the loaded
    positions (the reference's iteration 0)."""
    from .engine import Engine
    en = np.asarray(st.pbc_enabled, bool).reshape(-1, 3)
    pbc = np.where(en.any(axis=0) if st.n else np.zeros(3, bool), st.pbc, NO_PBC)
    eng = Engine(**engine_kw)
    eng.set_system(st.elem, st.pos, st.vel, pbc, st.tags.get("fixed", np.zeros(st.n, bool)).astype(np.uint8))
    if st.n and not (en == en[0]).all():
        eng.set_atom_pbc(np.where(en, st.pbc[None, :], NO_PBC), st.pos, st.pbc_count)
    else:
        eng.set_positions(st.pos, pbc_count=st.pbc_count)
    b = st.bath
    if b and b["geometry"] != 0:
        eng.set_thermal_bath(geometry=b["geometry"], To=b["To"], gamma=b["gamma"], z_min=b["z_min"],
                             d_boundary=b["d_boundary"], z_min_free=b["z_min_free"], center=b["center"],
                             radius=b["radius"])
        eng.set_bath_atoms(st.tags.get("thermal_bath_applicable", np.ones(st.n, bool)).astype(np.uint8))
    if clock:
        t = st.time
        eng.set_time(t["sim_time"], t["sim_time_final"], t["dt"])
    return eng


def update_state(st: State, eng, sim_time_advance: float = 0.0, dt: float | None = None) -> State:
    """A copy of `st` with the engine's positions, velocities, accelerations
    and PBC counters (and the time fields advanced)."""
    x, v, a, c = eng.state()
    time = dict(st.time)
    time["sim_time"] += sim_time_advance
    if dt is not None:
        time["dt_prev"], time["dt"] = time["dt"], dt
    return State(st.elem.copy(), x, v, a, st.pbc.copy(), c.astype(np.int32), st.pbc_enabled.copy(),
                 st.indices.copy(), {k: v.copy() for k, v in st.tags.items()},
                 None if st.bath is None else dict(st.bath), time, None if st.check is None else dict(st.check))
