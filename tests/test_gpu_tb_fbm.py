"""FBM (Born-Mayer Cu-Cu, Cu-Ar) and TightBinding (Cu second moment) on the
GPU against the oracle.

A periodic Cu block, an open Cu/Ag cluster with Ar atoms and close contacts
inside the Born-Mayer (< 1.0 A) and bridge (1.0-1.2 A) ranges, both members
together and beside Ackland, their neighbour lists, and a trajectory of an
Ar atom through a Cu block.  FP64 bar: energies per member 1e-12 relative,
forces 1e-10 of the largest force, trajectories 1e-10.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, force_rel_err

pytestmark = pytest.mark.gpu


def _cluster():
    rng = np.random.default_rng(12)
    base = S.fcc(64, 3, 3, 3, a=3.7, periodic=(False, False, False))
    p = base.pos + rng.normal(scale=0.2, size=base.pos.shape) + 10.0
    el = np.where(rng.random(base.n) < 0.2, 108, 64).astype(np.int32)
    el[0], el[4] = 64, 64
    p[1] = p[0] + [0.9, 0.0, 0.0]  # Born-Mayer range
    el[1] = 64
    p[5] = p[4] + [0.0, 1.1, 0.0]  # bridge range
    el[5] = 64
    ar = np.array([p[10] + [0.0, 0.0, 1.6], p[20] + [2.2, 0.0, 0.0], p[30] + [0.0, 3.3, 0.0]])
    return S.System(np.concatenate([el, [S.AR] * 3]).astype(np.int32), np.concatenate([p, ar]),
                    np.array([S.NO_PBC] * 3), name="Cu/Ag cluster + Ar")


def _pair(s, names, mask, **kw):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle
    return Oracle.from_system(s, pots=mask), Engine.from_system(s, potentials=names, **kw)


def _check(o, g, name):
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    print(f"{name}: E_oracle(eV)={np.asarray(eo) / EV}, E_gpu(eV)={np.asarray(eg) / EV}", flush=True)
    for x, y in zip(eg, eo):
        assert abs(x - y) <= 1e-12 * abs(y) + 1e-9 * EV
    fr = force_rel_err(gg, go)
    print(f"  force rel err {fr:.3e}", flush=True)
    assert fr <= 1e-10
    return eo


def test_tb_block_parity():
    from oracle.oracle import POT_TB
    s = S.jitter(S.fcc(64, 5, 5, 5), 0.12, 31)
    eo = _check(*_pair(s, ("TightBinding",), POT_TB), s.name)
    assert eo[0] < 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fbm_tb_cluster_with_close_contacts(precision):
    from oracle.oracle import POT_FBM, POT_TB
    s = _cluster()
    eo = _check(*_pair(s, ("FBM", "TightBinding"), POT_FBM | POT_TB, precision=precision), f"{s.name} ({precision})")
    assert eo[0] != 0.0 and eo[1] != 0.0


def test_metal_members_together():
    from oracle.oracle import POT_ACKLAND, POT_FBM, POT_TB
    s = _cluster()
    _check(*_pair(s, ("Ackland", "FBM", "TightBinding"), POT_ACKLAND | POT_FBM | POT_TB), s.name + " (3 members)")


@pytest.mark.parametrize("name,mask", [("FBM", 128), ("TightBinding", 256)])
def test_lists_match(name, mask):
    s = _cluster()
    o, g = _pair(s, (name,), mask)
    for i, (x, y) in enumerate(zip(o.neighbour_lists(0), g.neighbour_lists(0))):
        assert np.array_equal(x, y), i


def test_ar_through_copper_trajectory():
    from oracle.oracle import POT_FBM, POT_TB
    cu = S.jitter(S.fcc(64, 4, 4, 3, periodic=(True, True, False)), 0.05, 32)
    cu.vel = S.maxwell_velocities(cu, 300.0, stream=33)
    top = cu.pos[:, 2].max()
    pos = np.concatenate([cu.pos, [[cu.pbc[0] / 2 + 0.3, cu.pbc[1] / 2 + 0.2, top + 3.2]]])
    vel = np.concatenate([cu.vel, [[0.0, 0.0, -0.25]]])  # A/fs, ~130 eV Ar downwards
    s = S.System(np.concatenate([cu.elem, [S.AR]]).astype(np.int32), pos, cu.pbc, vel, name="Ar -> Cu")
    o, g = _pair(s, ("FBM", "TightBinding"), POT_FBM | POT_TB)
    ro = o.run(60)
    rg = g.run(60)
    print(f"dt range {ro['dt'].min():.3e}..{ro['dt'].max():.3e}", flush=True)
    assert np.array_equal(ro["dt"], rg["dt"]) or np.abs(rg["dt"] - ro["dt"]).max() <= 1e-10 * ro["dt"].max()
    assert np.abs(rg["epot"] - ro["epot"]).max() <= 1e-10 * np.abs(ro["epot"]).max()
    assert np.abs(rg["ekin"] - ro["ekin"]).max() <= 1e-10 * np.abs(ro["ekin"]).max()
