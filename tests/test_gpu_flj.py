"""FLJ (hydrocarbon-metal pairs) on the GPU against the oracle.

A Cu or Au fcc block with a hydrocarbon fragment above it, in the reference
chain FBZL -> FLJ -> REBO -> AIREBO -> ETors (common.h:46-68), plus single H
and C atoms placed inside the ZBL (1.0 A) and bridge (1.5 A) ranges of metal
atoms.  FP64 bar: energies per chain member 1e-12 relative, forces 1e-10 of
the largest force; FLJ's neighbour list as the oracle's; a short trajectory.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, force_rel_err

pytestmark = pytest.mark.gpu

POTS = ("FBZL", "FLJ", "REBO", "AIREBO", "ETors")


def _fcc(elem, a, n=3):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)], float)
    return ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3) * a), np.full(4 * n ** 3, elem, np.int32)


def _system(metal=64, a=3.615, contacts=True):
    mpos, mel = _fcc(metal, a)
    hc = S.hydrocarbon_cluster(60, 3)
    hpos = hc.pos - hc.pos.mean(axis=0)
    hpos[:, 2] += mpos[:, 2].max() + 2.6 - hpos[:, 2].min()
    hpos[:, :2] += mpos[:, :2].mean(axis=0)
    pos = [mpos, hpos]
    el = [mel, hc.elem]
    if contacts:
        top = mpos[np.argmax(mpos[:, 2])]
        side = mpos[np.argmin(mpos[:, 0])]
        pos.append(np.array([top + [0, 0, 1.0], side - [1.5, 0, 0]]))
        el.append(np.array([S.H, S.C], np.int32))
    p = np.concatenate(pos) + 20.0
    return S.System(np.concatenate(el).astype(np.int32), p, np.array([S.NO_PBC] * 3), name=f"metal{metal}+HC")


def _pair(s):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_FBZL, POT_FLJ, POT_REBO, POT_AIREBO, POT_ETORS
    o = Oracle.from_system(s, pots=POT_FBZL | POT_FLJ | POT_REBO | POT_AIREBO | POT_ETORS)
    g = Engine.from_system(s, potentials=POTS)
    return o, g


@pytest.mark.parametrize("metal,a", [(64, 3.615), (197, 4.078)])
def test_flj_energy_force_parity(metal, a):
    s = _system(metal, a)
    o, g = _pair(s)
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    print(f"{s.name}: E_oracle(eV)={eo / EV}, E_gpu(eV)={eg / EV}", flush=True)
    assert eo[1] != 0.0  # FLJ contributes
    for x, y in zip(eg, eo):
        assert abs(x - y) <= 1e-12 * abs(y) + 1e-9 * EV
    fr = force_rel_err(gg, go)
    print(f"  force rel err {fr:.3e}", flush=True)
    assert fr <= 1e-10


def test_flj_neighbour_list_and_trajectory():
    s = _system(64, 3.615, contacts=False)
    s.vel = S.maxwell_velocities(s, 300.0, stream=6)
    o, g = _pair(s)
    lo, lg = o.neighbour_lists(1), g.neighbour_lists(1)
    for i, (x, y) in enumerate(zip(lo, lg)):
        assert np.array_equal(x, y), i
    ro = o.run(30, fixed_dt=0.2e-15)
    rg = g.run(30, fixed_dt=0.2e-15)
    assert np.abs(rg["epot"] - ro["epot"]).max() <= 1e-10 * np.abs(ro["epot"]).max()
    assert np.abs(rg["ekin"] - ro["ekin"]).max() <= 1e-10 * np.abs(ro["ekin"]).max()
