"""Ackland EAM (Cu, Ag, Au) on the GPU against the oracle.

Periodic fcc blocks of each metal at 600 K-scale displacements, an open
three-metal cluster with pairs inside the ZBL (< 1.0 A) and bridge
(1.0-1.4 A) ranges, the whole mdtrajsim chain FBZL -> FLJ -> REBO -> AIREBO
-> ETors -> Ackland (common.h:46-68) on a hydrocarbon fragment over a Cu block
with an Ar ion, and a batch of blocks.  FP64 bar: energies per chain member
1e-12 relative, forces 1e-10 of the largest force; Ackland's neighbour lists
as the oracle's (Ackland alone, so the long list has Ackland's own range);
a short trajectory to 1e-10.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, force_rel_err

pytestmark = pytest.mark.gpu


def _block(el, n=5, jit=0.12, stream=1):
    return S.jitter(S.fcc(el, n, n, n), jit, stream)


def _cluster():
    rng = np.random.default_rng(9)
    base = S.fcc(64, 3, 3, 3, a=3.9, periodic=(False, False, False))
    p = base.pos + rng.normal(scale=0.2, size=base.pos.shape) + 10.0
    el = rng.choice([64, 108, 197], base.n).astype(np.int32)
    # close contacts: one pair at 0.9 A (ZBL) and one at 1.2 A (bridge)
    p[1] = p[0] + [0.9, 0.0, 0.0]
    p[5] = p[4] + [0.0, 1.2, 0.0]
    return S.System(el, p, np.array([S.NO_PBC] * 3), name="CuAgAu cluster")


def _pair(s, names, mask, **kw):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle
    return Oracle.from_system(s, pots=mask), Engine.from_system(s, potentials=names, **kw)


def _check(eo, go, eg, gg, name):
    print(f"{name}: E_oracle(eV)={np.asarray(eo) / EV}, E_gpu(eV)={np.asarray(eg) / EV}", flush=True)
    for x, y in zip(eg, eo):
        assert abs(x - y) <= 1e-12 * abs(y) + 1e-9 * EV
    fr = force_rel_err(gg, go)
    print(f"  force rel err {fr:.3e}", flush=True)
    assert fr <= 1e-10


@pytest.mark.parametrize("el", [64, 108, 197])
def test_ackland_block_parity(el):
    from oracle.oracle import POT_ACKLAND
    s = _block(el)
    o, g = _pair(s, ("Ackland",), POT_ACKLAND)
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    assert eo[0] < 0
    _check(eo, go, eg, gg, s.name)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ackland_cluster_with_short_contacts(precision):
    from oracle.oracle import POT_ACKLAND
    s = _cluster()
    o, g = _pair(s, ("Ackland",), POT_ACKLAND, precision=precision)
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    _check(eo, go, eg, gg, f"{s.name} ({precision})")


def test_full_mdtrajsim_chain():
    from oracle.oracle import POT_FBZL, POT_FLJ, POT_REBO, POT_AIREBO, POT_ETORS, POT_ACKLAND
    cu = S.jitter(S.fcc(64, 5, 5, 3, periodic=(True, True, False)), 0.08, 3)
    hc = S.hydrocarbon_cluster(50, 4)
    hp = hc.pos - hc.pos.mean(axis=0)
    hp[:, 2] += cu.pos[:, 2].max() + 2.6 - hp[:, 2].min()
    hp[:, :2] += cu.pbc[:2] / 2
    ar = np.array([[cu.pbc[0] / 2 + 1.0, cu.pbc[1] / 2, hp[:, 2].max() + 4.0]])
    pos = np.concatenate([cu.pos, hp, ar])
    el = np.concatenate([cu.elem, hc.elem, [S.AR]]).astype(np.int32)
    s = S.System(el, pos, cu.pbc, name="Cu + HC + Ar")
    names = ("FBZL", "FLJ", "REBO", "AIREBO", "ETors", "Ackland")
    o, g = _pair(s, names, POT_FBZL | POT_FLJ | POT_REBO | POT_AIREBO | POT_ETORS | POT_ACKLAND)
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    assert all(eo[k] != 0.0 for k in (0, 1, 2, 3, 5))
    _check(eo, go, eg, gg, s.name)


def test_ackland_lists_and_trajectory():
    from oracle.oracle import POT_ACKLAND
    s = _block(64, n=5, jit=0.05, stream=2)
    s.vel = S.maxwell_velocities(s, 600.0, stream=7)
    o, g = _pair(s, ("Ackland",), POT_ACKLAND)
    lo, lg = o.neighbour_lists(0), g.neighbour_lists(0)
    for i, (x, y) in enumerate(zip(lo, lg)):
        assert np.array_equal(x, y), i
    ro = o.run(40, fixed_dt=1e-15)
    rg = g.run(40, fixed_dt=1e-15)
    assert np.abs(rg["epot"] - ro["epot"]).max() <= 1e-10 * np.abs(ro["epot"]).max()
    assert np.abs(rg["ekin"] - ro["ekin"]).max() <= 1e-10 * np.abs(ro["ekin"]).max()


def test_ackland_batch_matches_single():
    from mdtk_amd import Engine
    mem = [_block(64, 5, stream=3), _block(197, 6, stream=4), _block(108, 5, stream=5)]
    bat = Engine.from_systems(mem, potentials=("Ackland",))
    eb, gb = bat.energy_grad()
    for k, s in enumerate(mem):
        one = Engine.from_system(s, potentials=("Ackland",))
        e1, g1 = one.energy_grad()
        a, b = bat.offsets[k], bat.offsets[k + 1]
        assert abs(eb[k][0] - e1[0]) <= 1e-12 * abs(e1[0])
        assert force_rel_err(gb[a:b], g1) <= 1e-12


def test_copper_slab_in_a_thermal_bath():
    """Ackland with SimLoop's box bath (To = 600 K on the slab's lower half and
    its borders) against the oracle, adaptive dt: energies and the energy
    taken from the bath."""
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_ACKLAND
    s = S.jitter(S.fcc(64, 5, 5, 4, periodic=(True, True, False)), 0.05, 9)
    s.vel = S.maxwell_velocities(s, 150.0, stream=10)
    zmid = (s.pos[:, 2].min() + s.pos[:, 2].max()) / 2
    kw = dict(To=600.0, gamma=5.0e13, z_min=zmid * S.ANG, d_boundary=2.0 * S.ANG, z_min_free=-5.0 * S.ANG)
    o = Oracle.from_system(s, pots=POT_ACKLAND)
    o.set_bath(2, kw["To"], kw["gamma"], kw["z_min"], kw["d_boundary"], kw["z_min_free"])
    g = Engine.from_system(s, potentials=("Ackland",))
    g.set_thermal_bath(geometry="box", **kw)
    ro, rg = o.run(60), g.run(60)
    assert np.abs(rg["epot"] - ro["epot"]).max() <= 1e-10 * np.abs(ro["epot"]).max()
    assert np.abs(rg["ekin"] - ro["ekin"]).max() <= 1e-10 * np.abs(ro["ekin"]).max()
    eb_o, eb_g = o.bath_energy(), g.bath_energy()[0]
    print(f"bath energy oracle {eb_o / EV:.6e} eV, gpu {eb_g / EV:.6e} eV", flush=True)
    assert eb_o > 0 and abs(eb_g - eb_o) <= 1e-9 * abs(eb_o)


@pytest.mark.parametrize("names,mask", [(("Ackland",), 64), (("Ackland", "TightBinding"), 64 | 256)])
def test_flying_atom_trajectory(names, mask):
    """A Cu atom flying at 0.1 A/fs into a 600 K copper surface, 200 steps of
    0.5 fs: its displacement rebuilds the metal rows on their own 1 A skin
    between long-list rebuilds (2.2 A skin), before and after the impact;
    energies stay with the oracle's."""
    b = S.jitter(S.fcc(64, 5, 5, 3, periodic=(True, True, False)), 0.05, 11)
    b.vel = S.maxwell_velocities(b, 600.0, stream=12)
    top = b.pos[:, 2].max()
    pos = np.concatenate([b.pos, [[b.pbc[0] / 2, b.pbc[1] / 2, top + 8.0]]])
    vel = np.concatenate([b.vel, [[0.0, 0.0, -0.1]]])  # A/fs
    s = S.System(np.concatenate([b.elem, [64]]).astype(np.int32), pos, b.pbc, vel, name="Cu -> Cu(001)")
    o, g = _pair(s, names, mask)
    ro, rg = o.run(200, fixed_dt=0.5e-15), g.run(200, fixed_dt=0.5e-15)
    assert np.abs(rg["epot"] - ro["epot"]).max() <= 1e-10 * np.abs(ro["epot"]).max()
    assert np.abs(rg["ekin"] - ro["ekin"]).max() <= 1e-10 * np.abs(ro["ekin"]).max()
    assert g.state()[0][-1, 2] < (top + 2.0) * S.ANG  # the atom reached the surface
