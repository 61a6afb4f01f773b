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
