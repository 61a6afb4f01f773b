"""ctypes front end of the oracle (the CPU restatement of the reference path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (mdtk_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

POT_FBZL, POT_BRENNER, POT_REBO, POT_AIREBO, POT_ETORS, POT_FLJ, POT_ACKLAND = 1, 2, 4, 8, 16, 32, 64
POT_FBM, POT_TB = 128, 256
NUM_PARAMS = 33  # REBO pair {Q, A, alpha, B1-3, beta1-3} x (CC, CH, HH), then LJ {eps, sigma} x 3
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ip, dp, dp, dp, ctypes.POINTER(ctypes.c_ubyte),
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_num_terms.argtypes = [ctypes.c_void_p]
        L.orc_term_kind.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_energy_grad.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_nl.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ip, ctypes.c_long]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp, dp, dp, dp]
        L.orc_get_state.argtypes = [ctypes.c_void_p, dp, dp, dp, ip]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_spline_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, dp]
        L.orc_set_alter_exact.argtypes = [ctypes.c_int]
        L.orc_set_params.argtypes = [ctypes.c_void_p, dp]
        L.orc_set_bath.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 5 + [dp, ctypes.c_double]
        L.orc_set_bath_atoms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ubyte)]
        L.orc_bath_energy.argtypes = [ctypes.c_void_p, dp]
        L.orc_get_params.argtypes = [ctypes.c_void_p, dp]
        L.orc_depos.argtypes = [dp, dp, ip, dp, dp, ip, dp]
        L.orc_apply_pbc.argtypes = [dp, dp, ip]
        L.orc_cos_theta.argtypes = [dp, ctypes.c_double, dp]
        L.orc_cos_dihedral.argtypes = [dp, ctypes.c_double, dp]
        L.orc_moved_too_much.argtypes = [ctypes.c_int, dp, ctypes.c_double, ip]
        L.orc_cij_path.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp, ip, ip]
        L.orc_virial.argtypes = [ctypes.c_void_p, dp]
        L.orc_set_atom_pbc.argtypes = [ctypes.c_void_p, dp, ip]
        _lib = L
    return _lib


def set_alter_exact(on: bool) -> None:
    """Test switch: zero derivative for the b* pair's constant r_alter."""
    lib().orc_set_alter_exact(int(bool(on)))


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


class OracleError(RuntimeError):
    pass


class Oracle:
    """One oracle simulation: a potential chain over one system (CGS arrays)."""

    def __init__(self, elem, xyz_cm, vel_cm_s, pbc_cm, fixed=None, pots=POT_REBO,
                 brenner_set=0, etors_compat=True, fbzl_r=(5.0e-8, 5.5e-8)):
        L = lib()
        self.n = int(len(elem))
        self._keep = [np.ascontiguousarray(elem, np.int32), np.ascontiguousarray(xyz_cm, np.float64),
                      None if vel_cm_s is None else np.ascontiguousarray(vel_cm_s, np.float64),
                      np.ascontiguousarray(pbc_cm, np.float64),
                      None if fixed is None else np.ascontiguousarray(fixed, np.uint8)]
        e, x, v, p, f = self._keep
        self.h = L.orc_create(self.n, _p(e, ctypes.c_int), _p(x), _p(v), _p(p), _p(f, ctypes.c_ubyte),
                              int(pots), int(brenner_set), int(bool(etors_compat)), fbzl_r[0], fbzl_r[1])
        if not self.h:
            raise OracleError(L.orc_last_error().decode())
        self.nterms = L.orc_num_terms(self.h)
        self.kinds = [L.orc_term_kind(self.h, i) for i in range(self.nterms)]

    @classmethod
    def from_system(cls, sys, **kw):
        e, x, v, p, f = sys.to_cgs()
        return cls(e, x, v, p, f, **kw)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(lib().orc_last_error().decode())

    def energy_grad(self):
        e = np.zeros(self.nterms)
        g = np.zeros((self.n, 3))
        self._chk(lib().orc_energy_grad(self.h, _p(e), _p(g)))
        return e, g

    def params(self):
        """The 33 differentiable parameters (REBO pair and LJ, CGS)."""
        th = np.zeros(NUM_PARAMS)
        self._chk(lib().orc_get_params(self.h, _p(th)))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float64)
        assert th.shape == (NUM_PARAMS,)
        self._chk(lib().orc_set_params(self.h, _p(th)))

    def set_bath(self, geometry: int, To: float, gamma: float = 1.0e13, z_min: float = 1e6 * 1e-8,
                 d_boundary: float = 0.0, z_min_free: float = -5e-8, center=(0.0, 0.0, 0.0),
                 radius: float = 1e6 * 1e-8):
        """Thermal bath (SimLoop.hpp:118-189): geometry 0 none, 1 everywhere,
        2 box, 3 sphere; To in K, gamma in 1/s, lengths in cm."""
        c = np.ascontiguousarray(center, np.float64)
        self._chk(lib().orc_set_bath(self.h, int(geometry), float(To), float(gamma), float(z_min),
                                     float(d_boundary), float(z_min_free), _p(c), float(radius)))

    def set_bath_atoms(self, apply):
        a = np.ascontiguousarray(apply, np.uint8)
        assert a.shape == (self.n,)
        self._chk(lib().orc_set_bath_atoms(self.h, _p(a, ctypes.c_ubyte)))

    def bath_energy(self) -> float:
        e = np.zeros(1)
        self._chk(lib().orc_bath_energy(self.h, _p(e)))
        return float(e[0])

    def neighbour_lists(self, term: int):
        counts = np.zeros(self.n, np.int32)
        cap = self.n * 1024
        idx = np.zeros(cap, np.int32)
        rc = lib().orc_nl(self.h, term, _p(counts, ctypes.c_int), _p(idx, ctypes.c_int), cap)
        if rc == -1:
            raise OracleError("neighbour list capacity")
        self._chk(rc)
        off = np.concatenate([[0], np.cumsum(counts)])
        return [idx[off[i]:off[i + 1]].copy() for i in range(self.n)]

    def run(self, nsteps: int, fixed_dt: float = 0.0):
        ep, ek, dt, nf = (np.zeros(nsteps) for _ in range(4))
        self._chk(lib().orc_run(self.h, nsteps, fixed_dt, _p(ep), _p(ek), _p(dt), _p(nf)))
        return dict(epot=ep, ekin=ek, dt=dt, netforce=nf)

    def virial(self):
        """Virial of one force call at the current positions: xx, yy, zz, xy,
        xz, yz in erg (sum over the gradient pushes of -r (x) v)."""
        w = np.zeros(6)
        self._chk(lib().orc_virial(self.h, _p(w)))
        return w

    def cij_path(self, i: int, j: int):
        """(weight, nodes) of the C_ij path stored for the LJ pair (i, j) at
        the last energy_grad (nodes[0] = i; empty when there is none)."""
        w = np.zeros(1)
        hops = np.zeros(1, np.int32)
        nodes = np.zeros(4, np.int32)
        self._chk(lib().orc_cij_path(self.h, int(i), int(j), _p(w), _p(hops, ctypes.c_int),
                                     _p(nodes, ctypes.c_int)))
        return float(w[0]), [int(q) for q in nodes[:hops[0] + 1]] if hops[0] else []

    def set_atom_pbc(self, pbc_cm, pbc_count=None):
        """Each atom's own PBC vector (n x 3, cm; NO_PBC for an open axis)
        and image counters, before the first call (Atom::PBC, PBC_count)."""
        p = np.ascontiguousarray(pbc_cm, np.float64).reshape(self.n, 3)
        c = None if pbc_count is None else np.ascontiguousarray(pbc_count, np.int32).reshape(self.n, 3)
        self._keep += [p, c]
        self._chk(lib().orc_set_atom_pbc(self.h, _p(p), _p(c, ctypes.c_int)))

    def state(self):
        x, v, a = (np.zeros((self.n, 3)) for _ in range(3))
        c = np.zeros((self.n, 3), np.int32)
        lib().orc_get_state(self.h, _p(x), _p(v), _p(a), _p(c, ctypes.c_int))
        return x, v, a, c


def spline_eval(table: int, x: float, y: float = 0.0, z: float = 0.0, which: int = 0) -> float:
    out = np.zeros(1)
    rc = lib().orc_spline_eval(table, x, y, z, which, _p(out))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return float(out[0])


def _chk(rc):
    if rc:
        raise OracleError(lib().orc_last_error().decode())


def depos(xa, La, ca, xb, Lb, cb):
    """depos(a, b) with per-atom PBC vectors and counters (cm)."""
    a = [np.ascontiguousarray(v, t) for v, t in ((xa, np.float64), (La, np.float64), (ca, np.int32),
                                                 (xb, np.float64), (Lb, np.float64), (cb, np.int32))]
    out = np.zeros(3)
    _chk(lib().orc_depos(_p(a[0]), _p(a[1]), _p(a[2], ctypes.c_int), _p(a[3]), _p(a[4]),
                         _p(a[5], ctypes.c_int), _p(out)))
    return out


def apply_pbc(xyz, L, pc):
    x = np.array(xyz, np.float64)
    c = np.array(pc, np.int32)
    Lb = np.ascontiguousarray(L, np.float64)
    _chk(lib().orc_apply_pbc(_p(x), _p(Lb), _p(c, ctypes.c_int)))
    return x, c


def cos_theta(xyz3, V=0.0):
    """(cos, grad[3,3]) of CosTheta(ij, ik, V) for atoms i, j, k."""
    x = np.ascontiguousarray(xyz3, np.float64).reshape(9)
    out = np.zeros(10)
    _chk(lib().orc_cos_theta(_p(x), float(V), _p(out)))
    return out[0], out[1:].reshape(3, 3)


def cos_dihedral(xyz4, V=0.0):
    """(cos, grad[4,3]) of CosDihedral(ij, ik, jl, V) for atoms i, j, k, l."""
    x = np.ascontiguousarray(xyz4, np.float64).reshape(12)
    out = np.zeros(13)
    _chk(lib().orc_cos_dihedral(_p(x), float(V), _p(out)))
    return out[0], out[1:].reshape(4, 3)


def moved_too_much(disp, rc):
    d = np.ascontiguousarray(disp, np.float64)
    out = np.zeros(1, np.int32)
    _chk(lib().orc_moved_too_much(len(d), _p(d), float(rc), _p(out, ctypes.c_int)))
    return bool(out[0])
