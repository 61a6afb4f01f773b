"""LJ pair rows against displacements (GPU).

The LJ pass reads per-atom pair rows: the long-list entries that lay within
the LJ cutoff plus a 1 A margin when the rows were built.  They are rebuilt
when the two largest displacements since then add up to more than the
margin.  Displacing atoms by prescribed amounts (handed over with dr, as a
driver that moves atoms itself does) checks both sides of that rule: below
the margin the rows are kept and the energies and forces still equal a fresh
evaluation at the new positions (the oracle for FP64, a fresh context for the
mixed form); above it the rows are rebuilt.  The moves change which pairs are
inside the LJ cutoff, so a kept row that missed a pair would show.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, force_rel_err, pe_small

pytestmark = pytest.mark.gpu

POTS = ("REBO", "AIREBO", "ETors")


def _reference(sys, precision):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO, POT_ETORS
    if precision == "fp64":
        return Oracle.from_system(sys, pots=POT_REBO | POT_AIREBO | POT_ETORS).energy_grad()
    return Engine.from_system(sys, potentials=POTS, precision=precision).energy_grad()


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pair_rows_follow_displacements(precision):
    from mdtk_amd import Engine
    base = pe_small(seed=5, jit=0.05)
    n = base.n
    eng = Engine.from_system(base, potentials=POTS, precision=precision)
    eng.energy_grad()
    c = eng.counters()
    assert c["pair_rows_built"] == n
    pairs = [c["lj_pairs"]]
    h = np.flatnonzero(base.elem == S.H)
    a, b = int(h[0]), int(h[len(h) // 2])
    dirs = _unit(S.normals(3 * n, 31).reshape(n, 3))
    d1 = 0.005 * dirs
    d1[a] = 0.45 * dirs[a]
    d1[b] = 0.45 * dirs[b]
    d2 = np.zeros((n, 3))
    d2[a] = 0.2 * dirs[a]  # a has now moved 0.65 A: 0.65 + 0.45 > 1 A
    x = base.pos.copy()
    for step, (d, built) in enumerate([(d1, n), (d2, 2 * n)]):
        x = x + d
        x = x - np.floor(x / base.pbc) * base.pbc  # back into the periodic box; dr carries the motion
        eng.set_positions(x * S.ANG, dr_cm=d * S.ANG)
        e, g = eng.energy_grad()
        c = eng.counters()
        pairs.append(c["lj_pairs"])
        moved = S.System(base.elem, x.copy(), base.pbc, name=f"pe moved {step}")
        er, gr = _reference(moved, precision)
        fr = force_rel_err(g, gr)
        print(f"{precision} move {step}: E={e.sum() / EV:.9f} eV, ref {er.sum() / EV:.9f} eV, "
              f"force rel err {fr:.2e}, LJ pairs {c['lj_pairs']}, rows built {c['pair_rows_built']}", flush=True)
        assert c["pair_rows_built"] == built
        for p, q in zip(e, er):
            assert abs(p - q) <= 1e-12 * abs(q) + 1e-9 * EV
        assert fr <= 1e-10
    assert len(set(pairs)) > 1  # the moves changed the set of LJ pairs
