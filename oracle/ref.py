"""ctypes front end of oracle/_ref/libmdtkref.so: the reference's own
GSL-free translation units (FBZL, FBM, NeighbourList, FGeneral's angle and
dihedral adjoints, Atom/AtomsArray, FProxy) compiled in place from the
reference tree by oracle/ref.mk.

TEST INFRASTRUCTURE ONLY: the CPU tests use it to check the oracle's
restatement against the reference code where that code builds here.  The
library exists only where the reference tree was present at build time;
available() says whether it was built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libmdtkref.so")
REF_TREE = "/root/reference"
_lib = None


def build() -> bool:
    """Builds the library when the reference tree is present; False otherwise."""
    if not os.path.isdir(REF_TREE):
        return False
    subprocess.run(["make", "-s", "-j8", "-f", os.path.join(HERE, "ref.mk"), "REF=" + REF_TREE], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return os.path.exists(LIB)


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_depos.argtypes = [dp, dp, ip, dp, dp, ip, dp]
        L.ref_apply_pbc.argtypes = [dp, dp, ip]
        L.ref_cos_theta.argtypes = [dp, ctypes.c_double, dp]
        L.ref_cos_dihedral.argtypes = [dp, ctypes.c_double, dp]
        L.ref_pair_chain.argtypes = [ctypes.c_int, ip, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     dp, dp]
        L.ref_neighbour_list.argtypes = [ctypes.c_int, ip, dp, dp, ctypes.c_double, ip, ctypes.c_int, ip, ip,
                                         ctypes.c_long]
        L.ref_moved_too_much.argtypes = [ctypes.c_int, dp, ctypes.c_double, ip]
        L.ref_remove_momentum.argtypes = [ctypes.c_int, ip, ctypes.POINTER(ctypes.c_ubyte), dp]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


def _chk(rc):
    if rc:
        raise RuntimeError(lib().ref_last_error().decode())


def depos(xa, La, ca, xb, Lb, cb):
    a = [np.ascontiguousarray(v, t) for v, t in ((xa, np.float64), (La, np.float64), (ca, np.int32),
                                                 (xb, np.float64), (Lb, np.float64), (cb, np.int32))]
    out = np.zeros(3)
    _chk(lib().ref_depos(_p(a[0]), _p(a[1]), _p(a[2], ctypes.c_int), _p(a[3]), _p(a[4]),
                         _p(a[5], ctypes.c_int), _p(out)))
    return out


def apply_pbc(xyz, L, pc):
    x = np.array(xyz, np.float64)
    c = np.array(pc, np.int32)
    Lb = np.ascontiguousarray(L, np.float64)
    _chk(lib().ref_apply_pbc(_p(x), _p(Lb), _p(c, ctypes.c_int)))
    return x, c


def cos_theta(xyz3, V=0.0):
    x = np.ascontiguousarray(xyz3, np.float64).reshape(9)
    out = np.zeros(10)
    _chk(lib().ref_cos_theta(_p(x), float(V), _p(out)))
    return out[0], out[1:].reshape(3, 3)


def cos_dihedral(xyz4, V=0.0):
    x = np.ascontiguousarray(xyz4, np.float64).reshape(12)
    out = np.zeros(13)
    _chk(lib().ref_cos_dihedral(_p(x), float(V), _p(out)))
    return out[0], out[1:].reshape(4, 3)


def pair_chain(elem, xyz, pbc, mask, r0=5.0e-8, r1=5.5e-8):
    """(energies per member, gradient) of FBZL (mask 1) and/or FBM (mask 128)."""
    e = np.ascontiguousarray(elem, np.int32)
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    L = np.ascontiguousarray(pbc, np.float64)
    nt = int(bool(mask & 1)) + int(bool(mask & 128))
    et = np.zeros(nt)
    g = np.zeros(len(e) * 3)
    _chk(lib().ref_pair_chain(len(e), _p(e, ctypes.c_int), _p(x), _p(L), int(mask), r0, r1, _p(et), _p(g)))
    return et, g.reshape(-1, 3)


def neighbour_list(elem, xyz, pbc, rc, handled):
    e = np.ascontiguousarray(elem, np.int32)
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    L = np.ascontiguousarray(pbc, np.float64)
    h = np.ascontiguousarray(handled, np.int32)
    n = len(e)
    counts = np.zeros(n, np.int32)
    cap = n * 1024
    idx = np.zeros(cap, np.int32)
    _chk(lib().ref_neighbour_list(n, _p(e, ctypes.c_int), _p(x), _p(L), float(rc), _p(h, ctypes.c_int), len(h),
                                  _p(counts, ctypes.c_int), _p(idx, ctypes.c_int), cap))
    off = np.concatenate([[0], np.cumsum(counts)])
    return [idx[off[i]:off[i + 1]].copy() for i in range(n)]


def moved_too_much(disp, rc):
    d = np.ascontiguousarray(disp, np.float64)
    out = np.zeros(1, np.int32)
    _chk(lib().ref_moved_too_much(len(d), _p(d), float(rc), _p(out, ctypes.c_int)))
    return bool(out[0])


def remove_momentum(elem, vel, fixed=None):
    e = np.ascontiguousarray(elem, np.int32)
    v = np.array(vel, np.float64).reshape(-1)
    f = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    _chk(lib().ref_remove_momentum(len(e), _p(e, ctypes.c_int), _p(f, ctypes.c_ubyte), _p(v)))
    return v.reshape(-1, 3)
