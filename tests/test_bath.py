"""Thermal bath of the oracle (SimLoop.cxx:280-388, SimLoop.hpp:118-189), CPU.

The bath adds the friction -V M gamma (1 - sqrt(To/T)) to the atoms inside
its geometry, T the kinetic temperature of those atoms at the start of the
step.  Pins, chosen so that a wrong sign, factor, geometry test or energy
bookkeeping fails one of them:

* free atoms (no force) in a bath everywhere with To = 0 lose kinetic energy
  as exp(-2 gamma t) (to O(gamma dt)), and heat up when To > T;
* for free atoms the energy handed to the bath equals the change of the
  kinetic energy exactly (dv with and without the bath differ by the bath's
  force only);
* no geometry, or atoms that opt out, leave the run equal bit for bit to
  the bath-free one (a box with zero-width strips is not empty: an atom that
  crosses a periodic face in a step is outside [0, L) when the kick tests it,
  as in the reference);
* the box (zMin, lateral boundary strips) and sphere (outside the radius)
  geometries act on exactly the atoms they contain.
"""
import numpy as np

from mdtk_amd import systems as S
from oracle.oracle import Oracle, POT_FBZL, POT_REBO, POT_AIREBO, POT_ETORS

EV = S.EV
ANG = S.ANG
KB = 8.617343e-5 * EV  # erg/K, mdtk/consts.cxx


def _gas(n=8, spacing=12.0, pbc=(S.NO_PBC, S.NO_PBC, S.NO_PBC), speed=0.02, stream=3):
    """n Ar atoms on a line 12 A apart (no pair within the ZBL range), random
    velocities of `speed` A/fs."""
    pos = np.zeros((n, 3))
    pos[:, 0] = 5.0 + spacing * np.arange(n)
    pos[:, 1] = 7.0
    pos[:, 2] = 6.0
    d = S.normals(3 * n, stream).reshape(n, 3)
    vel = speed * d / np.linalg.norm(d, axis=1, keepdims=True)
    return S.System(np.full(n, S.AR, np.int32), pos, np.array(pbc, float), vel, name="Ar gas")


def _ekin(o):
    v = o.state()[1]
    m = 39.948 * 1.6605655e-24
    return 0.5 * m * (v * v).sum(axis=1)


def test_free_atoms_cool_exponentially():
    g = _gas()
    o = Oracle.from_system(g, pots=POT_FBZL)
    gamma, dt, steps = 1.0e13, 0.1e-15, 200
    o.set_bath(1, 0.0, gamma)
    e0 = _ekin(o).sum()
    o.run(steps, fixed_dt=dt)
    ratio = _ekin(o).sum() / e0
    assert abs(ratio / np.exp(-2 * gamma * steps * dt) - 1.0) < 0.01, ratio


def test_free_gas_temperature_follows_closed_form():
    """Uniform friction gamma (1 - sqrt(To/T)) on a free gas gives
    dT/dt = -2 gamma (T - sqrt(To T)), i.e. sqrt(T) relaxes to sqrt(To) as
    exp(-gamma t): T(t) = (sqrt(To) + (sqrt(T0) - sqrt(To)) exp(-gamma t))^2.
    (Without the square root the curve would be To + (T0 - To) exp(-2 gamma t).)"""
    g = _gas()
    o = Oracle.from_system(g, pots=POT_FBZL)
    gamma, dt = 1.0e13, 0.1e-15
    t0 = _ekin(o).sum() / (1.5 * KB * g.n)
    to = t0 / 4.0
    o.set_bath(1, to, gamma)
    r = o.run(1000, fixed_dt=dt)
    for k in (99, 499, 999):
        t = (k + 1) * dt
        want = (np.sqrt(to) + (np.sqrt(t0) - np.sqrt(to)) * np.exp(-gamma * t)) ** 2
        got = r["ekin"][k] / (1.5 * KB * g.n)
        assert abs(got / want - 1.0) < 2e-3, (k, got / t0, want / t0)


def test_free_atoms_heat_below_target():
    g = _gas()
    o = Oracle.from_system(g, pots=POT_FBZL)
    t0 = _ekin(o).sum() / (1.5 * KB * g.n)
    o.set_bath(1, 4.0 * t0, 1.0e13)
    prev = _ekin(o).sum()
    for _ in range(5):
        o.run(20, fixed_dt=0.1e-15)
        cur = _ekin(o).sum()
        assert cur > prev
        prev = cur


def test_bath_energy_equals_kinetic_change_for_free_atoms():
    g = _gas()
    o = Oracle.from_system(g, pots=POT_FBZL)
    o.set_bath(1, 50.0, 3.0e13)
    e0 = _ekin(o).sum()
    r = o.run(150, fixed_dt=0.2e-15)
    eb = o.bath_energy()
    assert eb < 0  # the hot gas (~ 4e4 K) gives energy to a 50 K bath
    assert abs((r["ekin"][-1] - e0) - eb) <= 1e-12 * abs(eb)


def _pe_slab():
    s = S.pe_slab(2)
    s.vel = S.maxwell_velocities(s, 300.0, stream=5)
    return s


def test_empty_geometry_and_opted_out_atoms_change_nothing():
    pots = POT_REBO | POT_AIREBO | POT_ETORS
    s = _pe_slab()
    ref = Oracle.from_system(s, pots=pots)
    r0 = ref.run(20, fixed_dt=0.2e-15)
    box = Oracle.from_system(s, pots=pots)
    box.set_bath(0, 0.0, 1.0e14)  # no geometry
    r1 = box.run(20, fixed_dt=0.2e-15)
    out = Oracle.from_system(s, pots=pots)
    out.set_bath(1, 0.0, 1.0e14)
    out.set_bath_atoms(np.zeros(s.n, np.uint8))
    r2 = out.run(20, fixed_dt=0.2e-15)
    for r in (r1, r2):
        assert np.array_equal(r["epot"], r0["epot"]) and np.array_equal(r["ekin"], r0["ekin"])
    assert box.bath_energy() == 0.0 and out.bath_energy() == 0.0


def test_box_geometry_acts_on_strips_and_top_only():
    """Lateral PBC 40 x 40 A, open z: atoms within 3 A of a lateral face (and
    above zMinOfFreeZone) or above zMin are slowed; the others fly on."""
    n = 12
    rng_d = S.normals(3 * n, 9).reshape(n, 3)
    pos = np.array([[1.5, 20, 10], [38.5, 5, 20], [20, 1.0, 4], [28, 39.0, 14],     # strips
                    [20, 20, 30], [10, 30, 31],                                        # above zMin
                    [10, 10, 10], [30, 10, 4], [10, 30, 10], [30, 28, 22],             # interior
                    [1.0, 30, -8.0], [20, 20, 16]], float)                              # strip below free zone; interior
    d = pos[:, None, :] - pos[None, :, :]
    d[..., :2] -= 40.0 * np.round(d[..., :2] / 40.0)
    r = np.linalg.norm(d, axis=2) + 1e3 * np.eye(n)
    assert r.min() > 6.0  # no pair inside the ZBL range: free atoms
    vel = 0.02 * rng_d / np.linalg.norm(rng_d, axis=1, keepdims=True)
    g = S.System(np.full(n, S.AR, np.int32), pos, np.array([40.0, 40.0, S.NO_PBC]), vel, name="Ar box")
    o = Oracle.from_system(g, pots=POT_FBZL)
    o.set_bath(2, 0.0, 1.0e13, z_min=25.0 * ANG, d_boundary=3.0 * ANG, z_min_free=-5.0 * ANG)
    v0 = o.state()[1].copy()
    o.run(10, fixed_dt=0.1e-15)
    v1 = o.state()[1]
    slowed = np.linalg.norm(v1, axis=1) < np.linalg.norm(v0, axis=1) * (1 - 1e-4)
    assert slowed.tolist() == [True] * 6 + [False] * 6
    assert np.array_equal(v1[6:], v0[6:])


def test_sphere_geometry_acts_outside_the_radius():
    g = _gas(n=6, spacing=6.5)
    x = g.pos
    center = np.array([x[:, 0].mean(), 7.0, 6.0])
    o = Oracle.from_system(g, pots=POT_FBZL)
    o.set_bath(3, 0.0, 1.0e13, z_min_free=-5.0 * ANG, center=center * ANG, radius=10.0 * ANG)
    v0 = o.state()[1].copy()
    o.run(10, fixed_dt=0.1e-15)
    v1 = o.state()[1]
    outside = np.linalg.norm(x - center, axis=1) > 10.0
    assert outside.any() and (~outside).any()
    assert np.array_equal(v1[~outside], v0[~outside])
    assert (np.linalg.norm(v1[outside], axis=1) < np.linalg.norm(v0[outside], axis=1)).all()
