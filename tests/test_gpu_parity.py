"""GPU parity: the HIP engine against the oracle on the same inputs.

FP64 bar (stated by the task): per-potential energy |dE| <= 1e-12 |E| (plus a
1e-9 eV floor for terms that sum to ~0), forces max|dF|/max|F_ref| <= 1e-10;
neighbour lists identical (same sets, same ascending order).
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, force_rel_err, systems_for_parity, pe_small

pytestmark = pytest.mark.gpu

NAMES = {"FBZL": 1, "Brenner": 2, "REBO": 4, "AIREBO": 8, "ETors": 16}
E_REL = 1e-12
F_REL = 1e-10


def _pair(sys, pots, etors_compat=True, brenner_set=0):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle
    mask = 0
    for p in pots:
        mask |= NAMES[p]
    o = Oracle.from_system(sys, pots=mask, etors_compat=etors_compat, brenner_set=brenner_set)
    g = Engine.from_system(sys, potentials=pots, etors_compat=etors_compat, brenner_paramset=brenner_set)
    return o, g


def _check_energy_force(sys, pots, **kw):
    o, g = _pair(sys, pots, **kw)
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    print(f"{sys.name} n={sys.n} pots={pots} E_oracle(eV)={eo / EV} E_gpu(eV)={eg / EV}", flush=True)
    assert len(eo) == len(eg)
    for a, b in zip(eo, eg):
        assert abs(a - b) <= E_REL * abs(a) + 1e-9 * EV, (a / EV, b / EV)
    fr = force_rel_err(gg, go)
    print(f"  force rel err {fr:.3e}", flush=True)
    assert fr <= F_REL


@pytest.mark.parametrize("name", list(systems_for_parity().keys()))
def test_energy_force_parity(name):
    sys, kw = systems_for_parity()[name]
    _check_energy_force(sys, kw["pots"])


def test_etors_true_adjoint_parity():
    _check_energy_force(pe_small(), ("REBO", "AIREBO", "ETors"), etors_compat=False)


@pytest.mark.parametrize("ps", [0, 1])
def test_brenner_parity(ps):
    _check_energy_force(S.jitter(S.diamond(2, 2, 2), 0.05, 3), ("Brenner",), brenner_set=ps)
    _check_energy_force(S.hydrocarbon_cluster(48, 9), ("Brenner",), brenner_set=ps)


def test_diamond_special_case():
    """Perfect diamond: REBO -7.369926 eV/atom, Brenner-I -7.346135 eV/atom, zero forces."""
    from mdtk_amd import Engine
    d = S.diamond(2, 2, 2)
    e, g = Engine.from_system(d, potentials=("REBO",)).energy_grad()
    assert abs(e[0] / EV / d.n - (-7.369926)) < 5e-7
    assert np.abs(g).max() < 1e-15
    e, g = Engine.from_system(d, potentials=("Brenner",), brenner_paramset=0).energy_grad()
    assert abs(e[0] / EV / d.n - (-7.346135)) < 5e-7


def test_neighbour_lists_exact():
    sys = pe_small()
    o, g = _pair(sys, ("REBO", "AIREBO", "ETors"))
    for term in range(3):
        lo = o.neighbour_lists(term)
        lg = g.neighbour_lists(term)
        assert len(lo) == len(lg)
        for i, (a, b) in enumerate(zip(lo, lg)):
            assert np.array_equal(a, b), (term, i, a, b)


def test_zbl_projectile_parity():
    """PE slab (open z) with an Ar atom inside the ZBL range of the surface."""
    slab = S.polyethylene(2, 3, 3)
    pos = slab.pos.copy()
    pbc = slab.pbc.copy()
    pbc[2] = S.NO_PBC
    top = pos[:, 2].max()
    ar = np.array([[pbc[0] / 2 + 0.3, pbc[1] / 2 - 0.2, top + 2.2]])
    sys = S.System(np.concatenate([slab.elem, [S.AR]]).astype(np.int32), np.concatenate([pos, ar]), pbc,
                   name="pe+Ar")
    _check_energy_force(sys, ("FBZL", "REBO", "AIREBO", "ETors"))


def test_trajectory_parity():
    """50 fixed-dt Verlet steps: energies and final state against the oracle."""
    sys = pe_small(seed=3, jit=0.03)
    sys.vel = S.maxwell_velocities(sys, 300.0, stream=1)
    o, g = _pair(sys, ("REBO", "AIREBO", "ETors"))
    ro = o.run(50, fixed_dt=0.2e-15)
    rg = g.run(50, fixed_dt=0.2e-15)
    scale = np.abs(ro["epot"]).max()
    assert np.abs(ro["epot"] - rg["epot"]).max() <= 1e-10 * scale
    assert np.abs(ro["ekin"] - rg["ekin"]).max() <= 1e-9 * np.abs(ro["ekin"]).max()
    xo, vo, ao, co = o.state()
    xg, vg, ag, cg = g.state()
    assert np.array_equal(co, cg)
    assert np.abs(xo - xg).max() <= 1e-9 * S.ANG


def test_adaptive_dt_parity():
    """Reference dt rule 0.05 A / v_max clamped to [1e-20, 5e-16] s."""
    sys = S.jitter(S.diamond(2, 2, 2), 0.02, 4)
    sys.vel = S.maxwell_velocities(sys, 300.0, stream=2)
    o, g = _pair(sys, ("REBO",))
    ro = o.run(10)
    rg = g.run(10)
    assert np.allclose(ro["dt"], rg["dt"], rtol=1e-12, atol=0)
    assert abs(ro["epot"][-1] - rg["epot"][-1]) <= 1e-11 * abs(ro["epot"][-1])


def test_brenner_c1_trajectory():
    """The plain Brenner NVE config at its own size: periodic diamond, 8 x 8 x 8
    conventional cells (4096 C), 300 K Maxwell-Boltzmann velocities with zero
    net momentum, dt 0.5 fs: per-step energies and the state after 20 steps
    equal the oracle's."""
    s = S.diamond(8, 8, 8)
    assert s.n == 4096
    s.vel = S.maxwell_velocities(s, 300.0, stream=1)
    assert np.abs((s.vel * S.MASS_AMU[S.C]).sum(0)).max() <= 1e-12 * np.abs(s.vel).max() * s.n
    o, g = _pair(s, ("Brenner",))
    ro = o.run(20, fixed_dt=0.5e-15)
    rg = g.run(20, fixed_dt=0.5e-15)
    scale = np.abs(ro["epot"]).max()
    assert np.abs(ro["epot"] - rg["epot"]).max() <= 1e-10 * scale
    assert np.abs(ro["ekin"] - rg["ekin"]).max() <= 1e-9 * np.abs(ro["ekin"]).max()
    xo, vo = o.state()[:2]
    xg, vg = g.state()[:2]
    print(f"diamond 4096 Brenner: max |dx| {np.abs(xo - xg).max() / S.ANG:.2e} A", flush=True)
    assert np.abs(xo - xg).max() <= 1e-9 * S.ANG
    assert np.abs(vo - vg).max() <= 1e-8 * np.abs(vo).max()
