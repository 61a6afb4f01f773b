"""Shared system builders and comparison helpers for the tests."""
import numpy as np

from mdtk_amd import systems as S

EV = S.EV
ANG = S.ANG
EV_PER_A = EV / ANG  # erg/cm per eV/Angstrom


def force_rel_err(g_test, g_ref):
    """max_i |dF_i| / max_i |F_i^ref| (the parity measure of the north star)."""
    d = np.linalg.norm(np.asarray(g_test) - np.asarray(g_ref), axis=1).max()
    s = np.linalg.norm(np.asarray(g_ref), axis=1).max()
    return d / s


def pe_small(seed=7, jit=0.08):
    return S.jitter(S.polyethylene(2, 3, 5), jit, seed)


def systems_for_parity():
    """Small inputs the oracle evaluates in well under a second each."""
    out = {}
    out["ethane"] = (S.ethane(), dict(pots=("REBO",)))
    out["diamond64_jit"] = (S.jitter(S.diamond(2, 2, 2), 0.05, 11), dict(pots=("REBO",)))
    out["cluster60"] = (S.hydrocarbon_cluster(60, 3), dict(pots=("REBO", "AIREBO", "ETors")))
    out["pe360"] = (pe_small(), dict(pots=("REBO", "AIREBO", "ETors")))
    out["graphite_jit"] = (S.jitter(S.graphite(5, 3, 2), 0.04, 5), dict(pots=("REBO", "AIREBO", "ETors")))
    return out


def mixed_pbc_case(seed=3):
    """A bombardment frame whose atoms do not share one PBC vector: a PE slab
    periodic in x and y, and an Ar ion open on every axis (Atom::PBC =
    NO_PBC) 2.5 A above the surface and 0.4 A from the x = 0 face, so that
    the ion meets slab atoms across that face only through their unwrapped
    coordinates.  Slab atoms within 3 A of the far x face carry PBC_count.x
    = -1 (they sit at x - L for the ion), a few others +1.
    Returns (system, per-atom PBC (n x 3, cm), PBC_count (n x 3))."""
    s = S.bombard_target(288, seed, lateral=(8, 3), height=2.5)
    pos = s.pos.copy()
    pos[-1, 0] = 0.4
    pos[-1, 1] = 0.5 * s.pbc[1]
    sys_ = S.System(s.elem, pos, s.pbc.copy(), s.vel, s.fixed, "mixedpbc", dict(s.meta))
    n = sys_.n
    pbc = np.tile(np.where(sys_.pbc != S.NO_PBC, sys_.pbc * ANG, S.NO_PBC * ANG), (n, 1))
    pbc[-1, :] = S.NO_PBC * ANG
    cnt = np.zeros((n, 3), np.int32)
    x = sys_.pos[:-1, 0]
    cnt[:-1, 0] = np.where(x > sys_.pbc[0] - 3.0, -1, 0)
    cnt[5:-1:37, 0] = 1
    cnt[7:-1:41, 1] = -2
    return sys_, pbc, cnt
