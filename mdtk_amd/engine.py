"""ctypes binding of include/mdg.h and the Python face of the engine.

``Engine`` plays the role of the reference's ``SimLoop`` with its ``FProxy``
chain (mdtk/SimLoop.hpp, mdtk/potentials/FProxy.hpp): it owns one system, or
a batch of independent systems, on the GPU, returns per-potential energies and
Atom::grad from one force call (FProxy::operator(), FProxy.cxx:45-61) and
advances the velocity-Verlet loop (SimLoop::executeMain, SimLoop.cxx:252-445).
Units are the reference's CGS.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdg.so")

POT_FBZL, POT_BRENNER, POT_REBO, POT_AIREBO, POT_ETORS, POT_FLJ, POT_ACKLAND = 1, 2, 4, 8, 16, 32, 64
POT_FBM, POT_TB = 128, 256
NUM_PARAMS = 33  # REBO pair {Q, A, alpha, B1-3, beta1-3} x (CC, CH, HH), then LJ {eps, sigma} x (CC, CH, HH)
PARAM_NAMES = [f"{n}_{p}" for p in ("CC", "CH", "HH")
               for n in ("Q", "A", "alpha", "B1", "B2", "B3", "beta1", "beta2", "beta3")] + \
              [f"{n}_{p}" for p in ("CC", "CH", "HH") for n in ("eps", "sigma")]
PREC_FP64, PREC_MIXED = 0, 1
SYSTEMS_STOPPED = 2  # mdg_step_nve: some systems stopped on their own device errors, the rest ran
ERR_BITS = {1: "short-list overflow", 2: "long-list overflow", 4: "bond overflow", 8: "cos range",
            16: "negative coordination", 32: "queue overflow", 64: "spline domain", 128: "Ackland row overflow",
            256: "path overflow", 512: "type range"}  # MDG_ERR_*
_NAMES = {"FBZL": POT_FBZL, "FLJ": POT_FLJ, "Brenner": POT_BRENNER, "REBO": POT_REBO, "AIREBO": POT_AIREBO,
          "ETors": POT_ETORS, "Ackland": POT_ACKLAND, "FBM": POT_FBM, "TightBinding": POT_TB}


class EngineError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [("pot_mask", ctypes.c_int), ("brenner_paramset", ctypes.c_int), ("etors_compat", ctypes.c_int),
                ("precision", ctypes.c_int), ("fbzl_r0", ctypes.c_double), ("fbzl_r1", ctypes.c_double),
                ("max_short_nb", ctypes.c_int), ("max_long_nb", ctypes.c_int)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libmdg.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"{path} is missing: build it with `python -m mdtk_amd.build`")
    L = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    L.mdg_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(vp)]
    L.mdg_destroy.argtypes = [vp]
    L.mdg_last_error.argtypes = [vp]
    L.mdg_last_error.restype = ctypes.c_char_p
    L.mdg_set_system.argtypes = [vp, ctypes.c_int, ip, dp, dp, dp, ctypes.POINTER(ctypes.c_uint8)]
    L.mdg_set_systems.argtypes = [vp, ctypes.c_int, ip, ip, dp, dp, dp, ctypes.POINTER(ctypes.c_uint8)]
    L.mdg_num_systems.argtypes = [vp]
    L.mdg_set_positions.argtypes = [vp, dp, ip, dp]
    L.mdg_set_atom_pbc.argtypes = [vp, dp, dp, ip]
    L.mdg_set_graph_block.argtypes = [vp, ctypes.c_int]
    L.mdg_work_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.mdg_bstar_bounds.argtypes = [vp, dp]
    L.mdg_get_params.argtypes = [vp, dp]
    L.mdg_set_params.argtypes = [vp, dp]
    L.mdg_energy_grad_params.argtypes = [vp, dp, dp, dp]
    L.mdg_num_terms.argtypes = [vp]
    L.mdg_term_kind.argtypes = [vp, ctypes.c_int]
    L.mdg_energy_grad.argtypes = [vp, dp, dp]
    L.mdg_nl_export.argtypes = [vp, ctypes.c_int, ip, ip, ctypes.c_int64]
    L.mdg_step_nve.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, dp, dp, dp]
    L.mdg_get_state.argtypes = [vp, dp, dp, dp, ip]
    L.mdg_set_stream.argtypes = [vp, vp]
    L.mdg_profile.argtypes = [vp, ctypes.c_int]
    L.mdg_kernel_time.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int64)]
    L.mdg_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.mdg_event_record.argtypes = [vp, ctypes.c_int]
    L.mdg_event_elapsed.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.mdg_synchronize.argtypes = [vp]
    L.mdg_system_errors.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
    L.mdg_set_tune.argtypes = [vp, dp]
    L.mdg_set_accelerations.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_uint8)]
    L.mdg_set_iteration.argtypes = [vp, ctypes.c_int, ctypes.c_int64]
    L.mdg_get_iteration.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.mdg_set_virial.argtypes = [vp, ctypes.c_int]
    L.mdg_virial.argtypes = [vp, dp]
    L.mdg_step_virial.argtypes = [vp, dp]
    L.mdg_energy_grad_noalloc.argtypes = [ctypes.POINTER(_Config), ctypes.c_int, ip, dp, dp, dp, dp]
    L.mdg_autotune.argtypes = [vp, ctypes.c_int, dp]
    L.mdg_set_thermal_bath.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Bath)]
    L.mdg_set_bath_atoms.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8)]
    L.mdg_bath_energy.argtypes = [vp, dp]
    L.mdg_set_time.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.mdg_get_time.argtypes = [vp, dp]
    L.mdg_set_trajectory_output.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int]
    L.mdg_trajectory_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.mdg_take_trajectory.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p]
    _lib = L
    return L


# mdg_frame and mdg_atom_record of mdg.h
FRAME_DTYPE = np.dtype([("kind", "<i4"), ("system", "<i4"), ("iteration", "<i8"), ("sim_time", "<f8"),
                        ("dt", "<f8"), ("first", "<i8"), ("count", "<i8")])
RECORD_DTYPE = np.dtype([("x", "<f8", 3), ("v", "<f8", 3), ("pbc_count", "<i4", 3), ("atom", "<i4")])
FRAME_SNAPSHOT, FRAME_XVA = 1, 2


class _Bath(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int), ("To", ctypes.c_double), ("gamma", ctypes.c_double),
                ("z_min", ctypes.c_double), ("d_boundary", ctypes.c_double), ("z_min_free", ctypes.c_double),
                ("center", ctypes.c_double * 3), ("radius", ctypes.c_double)]


BATH_GEOMETRIES = {"none": 0, "universe": 1, "box": 2, "sphere": 3}


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def _config(potentials, precision, etors_compat, brenner_paramset, fbzl_r):
    mask = 0
    for p in potentials:
        mask |= _NAMES[p]
    return _Config(mask, int(brenner_paramset), int(bool(etors_compat)),
                   PREC_FP64 if precision == "fp64" else PREC_MIXED, float(fbzl_r[0]), float(fbzl_r[1]), 0, 0)


def energy_grad_noalloc(sys, potentials=("REBO", "AIREBO", "ETors"), precision: str = "fp64",
                        etors_compat: bool = True, brenner_paramset: int = 0, fbzl_r=(5.0e-8, 5.5e-8)):
    """One force call through mdg_energy_grad_noalloc (no device memory is
    allocated; at most 4096 atoms): (per-member energies in erg, dE/dx in
    erg/cm) of the system at its positions."""
    L = load_library()
    cfg = _config(potentials, precision, etors_compat, brenner_paramset, fbzl_r)
    e, x, _, p, _ = sys.to_cgs()
    nterms = sum(1 for k in _NAMES.values() if cfg.pot_mask & k)
    et = np.zeros(nterms)
    g = np.zeros((len(e), 3))
    rc = L.mdg_energy_grad_noalloc(ctypes.byref(cfg), len(e), _p(e, ctypes.c_int32), _p(x), _p(p), _p(et), _p(g))
    if rc != 0:
        raise EngineError(L.mdg_last_error(None).decode())
    return et, g


class Engine:
    """One GPU context: a potential chain over one system.

    potentials: names from {"FBZL", "FLJ", "Brenner", "REBO", "AIREBO", "ETors",
    "Ackland", "FBM", "TightBinding"}; the chain order is the reference's
    (applications/common.h:46-68), FBM and TightBinding after it.
    """

    def __init__(self, potentials=("REBO", "AIREBO", "ETors"), precision: str = "fp64",
                 etors_compat: bool = True, brenner_paramset: int = 0, fbzl_r=(5.0e-8, 5.5e-8)):
        L = load_library()
        cfg = _config(potentials, precision, etors_compat, brenner_paramset, fbzl_r)
        h = ctypes.c_void_p()
        if L.mdg_create(ctypes.byref(cfg), ctypes.byref(h)) != 0:
            raise EngineError(L.mdg_last_error(None).decode())
        self._h = h
        self.n = 0
        self.nsys = 0
        self.offsets = np.zeros(1, np.int64)
        self.nterms = L.mdg_num_terms(h)
        self.kinds = [L.mdg_term_kind(h, i) for i in range(self.nterms)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.mdg_destroy(h)
            self._h = None

    def _chk(self, rc):
        if rc != 0:
            raise EngineError(_lib.mdg_last_error(self._h).decode())

    def set_system(self, elem, xyz_cm, vel_cm_s, pbc_cm, fixed=None):
        e = np.ascontiguousarray(elem, np.int32)
        x = np.ascontiguousarray(xyz_cm, np.float64).reshape(-1)
        v = None if vel_cm_s is None else np.ascontiguousarray(vel_cm_s, np.float64).reshape(-1)
        p = np.ascontiguousarray(pbc_cm, np.float64)
        f = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        self._chk(_lib.mdg_set_system(self._h, int(e.shape[0]), _p(e, ctypes.c_int32), _p(x), _p(v), _p(p),
                                      _p(f, ctypes.c_uint8)))
        self.n = int(e.shape[0])
        self.nsys = 1
        self.offsets = np.array([0, self.n], np.int64)
        return self

    def set_systems(self, parts):
        """A batch: parts is a sequence of (elem, xyz_cm, vel_cm_s, pbc_cm,
        fixed) tuples, one per independent system."""
        counts = np.array([len(q[0]) for q in parts], np.int32)
        e = np.ascontiguousarray(np.concatenate([np.asarray(q[0], np.int32) for q in parts]))
        x = np.ascontiguousarray(np.concatenate([np.asarray(q[1], np.float64).reshape(-1) for q in parts]))
        vs = [q[2] for q in parts]
        v = None
        if any(a is not None for a in vs):
            v = np.ascontiguousarray(np.concatenate(
                [np.zeros(3 * len(q[0])) if q[2] is None else np.asarray(q[2], np.float64).reshape(-1)
                 for q in parts]))
        p = np.ascontiguousarray(np.concatenate([np.asarray(q[3], np.float64).reshape(3) for q in parts]))
        fs = [q[4] if len(q) > 4 else None for q in parts]
        f = None
        if any(a is not None for a in fs):
            f = np.ascontiguousarray(np.concatenate(
                [np.zeros(len(q[0]), np.uint8) if fq is None else np.asarray(fq, np.uint8)
                 for q, fq in zip(parts, fs)]))
        self._chk(_lib.mdg_set_systems(self._h, len(parts), _p(counts, ctypes.c_int32), _p(e, ctypes.c_int32),
                                       _p(x), _p(v), _p(p), _p(f, ctypes.c_uint8)))
        self.n = int(e.shape[0])
        self.nsys = len(parts)
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        return self

    @classmethod
    def from_system(cls, sys, **kw):
        eng = cls(**kw)
        return eng.set_system(*sys.to_cgs())

    @classmethod
    def from_systems(cls, systems, **kw):
        eng = cls(**kw)
        return eng.set_systems([s.to_cgs() for s in systems])

    def energy_grad(self):
        """Energies per chain member (shape (nterms,), or (nsys, nterms) for
        a batch) in erg and dE/dx (n, 3) in erg/cm."""
        e = np.zeros(max(self.nsys, 1) * self.nterms)
        g = np.zeros((self.n, 3))
        self._chk(_lib.mdg_energy_grad(self._h, _p(e), _p(g)))
        return (e if self.nsys <= 1 else e.reshape(self.nsys, self.nterms)), g

    def set_virial(self, on: bool = True):
        """Every later force call also reduces the per-system virial."""
        self._chk(_lib.mdg_set_virial(self._h, int(bool(on))))
        self._virial = bool(on)

    def virial(self):
        """Virial of the last force call, (6,) or (nsys, 6): xx, yy, zz, xy,
        xz, yz in erg."""
        m = max(self.nsys, 1)
        w = np.zeros(6 * m)
        self._chk(_lib.mdg_virial(self._h, _p(w)))
        return w if self.nsys <= 1 else w.reshape(m, 6)

    def set_thermal_bath(self, geometry="box", To=300.0, gamma=1.0e13, z_min=1.0e6 * 1e-8, d_boundary=0.0,
                         z_min_free=-5.0e-8, center=(0.0, 0.0, 0.0), radius=1.0e6 * 1e-8, system=-1):
        """Thermal bath of SimLoop (SimLoop.hpp:118-189) for one system or all
        (system=-1): geometry "none", "universe", "box" or "sphere"; To in K,
        gamma in 1/s, lengths in cm (defaults are the reference's)."""
        tb = _Bath(BATH_GEOMETRIES[geometry] if isinstance(geometry, str) else int(geometry), float(To), float(gamma),
                   float(z_min), float(d_boundary), float(z_min_free), (ctypes.c_double * 3)(*map(float, center)),
                   float(radius))
        self._chk(_lib.mdg_set_thermal_bath(self._h, int(system), ctypes.byref(tb)))

    def set_bath_atoms(self, apply):
        """Atom::apply_ThermalBath per atom (None: all)."""
        a = None if apply is None else np.ascontiguousarray(apply, np.uint8)
        self._chk(_lib.mdg_set_bath_atoms(self._h, _p(a, ctypes.c_uint8)))

    def bath_energy(self):
        """Energy each system took from its bath so far (erg, per system)."""
        out = np.zeros(self.nsys)
        self._chk(_lib.mdg_bath_energy(self._h, _p(out)))
        return out

    def set_time(self, sim_time: float = 0.0, sim_time_final: float = 1e300, dt: float = 1e-20, system: int = -1):
        """SimLoop's simTime, simTimeFinal and dt (s) of one system or all (-1);
        a system past its final time stands still in run() (dt_used 0)."""
        self._chk(_lib.mdg_set_time(self._h, system, sim_time, sim_time_final, dt))

    def set_trajectory_output(self, snapshot_atoms=None, xva: bool = False):
        """Capture SnapshotList records of the flagged atoms (mdtrajsim's
        cadence) and/or XVA frames of whole systems inside run(); None and
        False turn it off."""
        sel = None if snapshot_atoms is None else np.ascontiguousarray(snapshot_atoms, np.uint8)
        self._chk(_lib.mdg_set_trajectory_output(self._h, _p(sel, ctypes.c_uint8), int(bool(xva))))

    def take_trajectory(self):
        """The frames (structured array, FRAME_DTYPE) and atom records
        (RECORD_DTYPE) captured since the last call."""
        nf, nr = ctypes.c_int64(0), ctypes.c_int64(0)
        self._chk(_lib.mdg_trajectory_size(self._h, ctypes.byref(nf), ctypes.byref(nr)))
        fr = np.zeros(nf.value, FRAME_DTYPE)
        rec = np.zeros(nr.value, RECORD_DTYPE)
        self._chk(_lib.mdg_take_trajectory(self._h, fr.ctypes.data if nf.value else None,
                                           rec.ctypes.data if nr.value else None))
        return fr, rec

    def time(self):
        """Per system (simTime, simTimeFinal, next dt), shape (nsys, 3)."""
        out = np.zeros(3 * max(self.nsys, 1))
        self._chk(_lib.mdg_get_time(self._h, _p(out)))
        return out.reshape(-1, 3)

    def params(self):
        """The NUM_PARAMS differentiable parameters (CGS), see PARAM_NAMES."""
        th = np.zeros(NUM_PARAMS)
        self._chk(_lib.mdg_get_params(self._h, _p(th)))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float64).reshape(NUM_PARAMS)
        self._chk(_lib.mdg_set_params(self._h, _p(th)))

    def energy_grad_params(self):
        """energy_grad plus dE_s/dtheta per system: (nsys, NUM_PARAMS), or
        (NUM_PARAMS,) for one system.  FP64 contexts only."""
        m = max(self.nsys, 1)
        e = np.zeros(m * self.nterms)
        g = np.zeros((self.n, 3))
        d = np.zeros(m * NUM_PARAMS)
        self._chk(_lib.mdg_energy_grad_params(self._h, _p(e), _p(g), _p(d)))
        if m == 1:
            return e, g, d
        return e.reshape(m, self.nterms), g, d.reshape(m, NUM_PARAMS)

    def set_positions(self, xyz_cm, pbc_count=None, dr_cm=None):
        x = np.ascontiguousarray(xyz_cm, np.float64).reshape(-1)
        c = None if pbc_count is None else np.ascontiguousarray(pbc_count, np.int32).reshape(-1)
        d = None if dr_cm is None else np.ascontiguousarray(dr_cm, np.float64).reshape(-1)
        self._chk(_lib.mdg_set_positions(self._h, _p(x), _p(c, ctypes.c_int32), _p(d)))

    def set_atom_pbc(self, pbc_cm, xyz_cm, pbc_count=None):
        """Each atom's own PBC vector (n x 3, cm; NO_PBC for an open axis),
        with the positions and image counters the caller's atoms hold
        (Atom::PBC, Atom::PBC_count); every list is rebuilt."""
        p = np.ascontiguousarray(pbc_cm, np.float64).reshape(-1)
        x = np.ascontiguousarray(xyz_cm, np.float64).reshape(-1)
        c = None if pbc_count is None else np.ascontiguousarray(pbc_count, np.int32).reshape(-1)
        self._chk(_lib.mdg_set_atom_pbc(self._h, _p(p), _p(x), _p(c, ctypes.c_int32)))

    def neighbour_lists(self, term: int):
        counts = np.zeros(self.n, np.int32)
        cap = self.n * 1024
        idx = np.zeros(cap, np.int32)
        self._chk(_lib.mdg_nl_export(self._h, int(term), _p(counts, ctypes.c_int32), _p(idx, ctypes.c_int32), cap))
        off = np.concatenate([[0], np.cumsum(counts)])
        return [idx[off[i]:off[i + 1]].copy() for i in range(self.n)]

    def run(self, nsteps: int, fixed_dt: float = 0.0, allow_stopped: bool = False):
        """nsteps Verlet iterations; per-step arrays of shape (nsteps,), or
        (nsteps, nsys) for a batch.  A system whose own terms raise a device
        error stands still from then on while the others run; with
        allow_stopped the outputs come back with out["stopped"] (the error
        bits per system, 0 = ran every step), otherwise EngineError."""
        m = max(self.nsys, 1)
        ep, ek, dt, nf = (np.zeros(nsteps * m) for _ in range(4))
        rc = _lib.mdg_step_nve(self._h, int(nsteps), float(fixed_dt), _p(ep), _p(ek), _p(dt), _p(nf))
        stopped = None
        if rc == SYSTEMS_STOPPED and allow_stopped:
            stopped = self.system_errors()
        else:
            self._chk(rc)
        out = dict(epot=ep, ekin=ek, dt=dt, netforce=nf)
        if m > 1:
            out = {k: v.reshape(nsteps, m) for k, v in out.items()}
        if getattr(self, "_virial", False):
            w = np.zeros(nsteps * m * 6)
            self._chk(_lib.mdg_step_virial(self._h, _p(w)))
            out["virial"] = w.reshape(nsteps, 6) if m == 1 else w.reshape(nsteps, m, 6)
        if stopped is not None:
            out["stopped"] = stopped
        return out

    def system_errors(self):
        """Device error bits of each system since the last query (0 = none),
        cleared by the query."""
        out = np.zeros(max(self.nsys, 1), np.uint32)
        self._chk(_lib.mdg_system_errors(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    def set_accelerations(self, acc, keep=None):
        """Accelerations (cm/s^2) of loaded states: the first run() starts the
        systems with keep[s] (None: all) from them instead of from the forces
        at their positions."""
        a = np.ascontiguousarray(acc, np.float64).reshape(self.n, 3)
        k = None if keep is None else np.ascontiguousarray(keep, np.uint8).reshape(max(self.nsys, 1))
        self._chk(_lib.mdg_set_accelerations(self._h, _p(a), _p(k, ctypes.c_uint8)))

    def set_iteration(self, iteration: int, system: int = -1):
        """SimLoop::iteration of one system or all (-1)."""
        self._chk(_lib.mdg_set_iteration(self._h, int(system), int(iteration)))

    def iterations(self):
        """Per system: SimLoop::iteration (steps done since 0 or the value set)."""
        out = np.zeros(max(self.nsys, 1), np.int64)
        self._chk(_lib.mdg_get_iteration(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def state(self):
        x, v, a = (np.zeros((self.n, 3)) for _ in range(3))
        c = np.zeros((self.n, 3), np.int32)
        self._chk(_lib.mdg_get_state(self._h, _p(x), _p(v), _p(a), _p(c, ctypes.c_int32)))
        return x, v, a, c

    def set_stream(self, stream_handle: int):
        self._chk(_lib.mdg_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def profile(self, on: bool):
        self._chk(_lib.mdg_profile(self._h, int(bool(on))))

    def kernel_time(self, name: str):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        self._chk(_lib.mdg_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def record(self, slot: int):
        """Records HIP event `slot` (0-7) on the engine's stream."""
        self._chk(_lib.mdg_event_record(self._h, int(slot)))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = ctypes.c_double()
        self._chk(_lib.mdg_event_elapsed(self._h, int(a), int(b), ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps: int = 3) -> dict:
        """Picks the force kernels' launch shapes by timing candidates on this
        engine's systems; returns the choices and the tuned kernels' ms per
        force call before and after."""
        r = np.zeros(10)
        self._chk(_lib.mdg_autotune(self._h, int(reps), _p(r)))
        keys = ["occ_bonds", "grid_bonds", "occ_bstar", "grid_bstar", "grid_lj", "grid_tors", "ms_before",
                "ms_after", "lj_shape", "xcd_order"]
        return {k: (float(v) if k.startswith("ms") else int(v)) for k, v in zip(keys, r)}

    def set_tune(self, tune: dict):
        """Launch shapes as autotune() returns them (e.g. saved from an earlier
        run on the same workload)."""
        keys = ["occ_bonds", "grid_bonds", "occ_bstar", "grid_bstar", "grid_lj", "grid_tors", "ms_before",
                "ms_after", "lj_shape", "xcd_order"]
        r = np.array([float(tune.get(k, 0.0)) for k in keys])
        self._chk(_lib.mdg_set_tune(self._h, _p(r)))

    def work_counts(self) -> dict:
        """Bond-order work of the last force call: REBO sigma-pi triplets, the b*
        bound pass's triplets and the b* pairs that took the full bond order."""
        w = np.zeros(3, np.int64)
        self._chk(_lib.mdg_work_counts(self._h, _p(w, ctypes.c_int64)))
        return dict(rebo_triplets=int(w[0]), bstar_triplets=int(w[1]), bstar_full=int(w[2]))

    def bstar_bounds(self) -> dict:
        """The table bounds of the per-atom b* test (G_C, G_H, P_CC, P_CH lower bounds)."""
        b = np.zeros(4)
        self._chk(_lib.mdg_bstar_bounds(self._h, _p(b)))
        return dict(g_c=b[0], g_h=b[1], p_cc=b[2], p_ch=b[3])

    def set_graph_block(self, steps: int):
        """Steps per graph launch in run() (default 32; results do not depend on it)."""
        self._chk(_lib.mdg_set_graph_block(self._h, int(steps)))

    def synchronize(self):
        self._chk(_lib.mdg_synchronize(self._h))

    def counters(self):
        out = (ctypes.c_int64 * 8)()
        self._chk(_lib.mdg_counters(self._h, out))
        keys = ["bonds", "lj_pairs", "bstar", "quads", "short_entries", "long_entries", "rebuilds",
                "pair_rows_built"]
        return dict(zip(keys, list(out)))
