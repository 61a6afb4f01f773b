"""Device errors of one system and the dt rule's hasNeighbors.

A batch is many independent trajectories; the reference runs one at a time
and a REQUIRE ends it (SimLoop.cxx:176-205).  Here a system whose own terms
raise a device error (a hydrogen 'hedgehog': one carbon with 18 hydrogens
inside its REBO range, more than the 16 bonds an atom may hold) stops, and
every other system of the batch runs every step exactly as it runs alone.

FProxy::hasNeighbors (FProxy.hpp:99-104, used by the dt rule at
SimLoop.cxx:394-398) asks each member's own neighbour list, and each list
holds only the elements that member handles (NeighbourList.cxx:90-91): a fast
Ar next to a hydrocarbon does not set dt under REBO + AIREBO + ETors (no
member handles Ar), and does under FBZL (which handles every element).
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import ANG

pytestmark = pytest.mark.gpu

CHAIN = ("REBO", "AIREBO", "ETors")
BOND_OVERFLOW = 4  # MDG_ERR_BOND_OVERFLOW


def _hedgehog(nh=18, r=1.12):
    k = np.arange(nh) + 0.5
    th = np.arccos(1 - 2 * k / nh)
    ph = np.pi * (1 + 5 ** 0.5) * k
    h = r * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    pos = np.vstack([[0.0, 0.0, 0.0], h]) + 30.0
    el = np.array([S.C] + [S.H] * nh, np.int32)
    return S.System(el, pos, np.array([S.NO_PBC] * 3), name="hedgehog")


def _split(x, counts):
    off = np.concatenate([[0], np.cumsum(counts)])
    return [x[off[k]:off[k + 1]] for k in range(len(counts))]


def test_stopped_system_leaves_the_batch_running():
    from mdtk_amd import Engine, EngineError
    pe = S.jitter(S.polyethylene(2, 3, 5), 0.05, 3)
    pe.vel = S.maxwell_velocities(pe, 300.0, stream=4)
    cl = S.hydrocarbon_cluster(60, 7)
    systems = [pe, _hedgehog(), cl]
    nsteps, dt = 4, 0.1e-15
    b = Engine.from_systems(systems, potentials=CHAIN, etors_compat=False)
    r = b.run(nsteps, fixed_dt=dt, allow_stopped=True)
    print("stopped:", r["stopped"], flush=True)
    assert r["stopped"][0] == 0 and r["stopped"][2] == 0
    assert r["stopped"][1] & BOND_OVERFLOW
    t = b.time()
    assert np.allclose(t[[0, 2], 0], nsteps * dt, rtol=1e-12)
    assert t[1, 0] == pytest.approx(dt, rel=1e-12)  # the step that raised it ran, then it stood still
    assert np.all(r["dt"][1:, 1] == 0.0) and np.all(r["dt"][:, [0, 2]] == dt)
    xb, vb, _, _ = b.state()
    xs, vs = _split(xb, [s.n for s in systems]), _split(vb, [s.n for s in systems])
    for k in (0, 2):
        e = Engine.from_system(systems[k], potentials=CHAIN, etors_compat=False)
        r1 = e.run(nsteps, fixed_dt=dt)
        x1, v1, _, _ = e.state()
        assert np.array_equal(xs[k], x1) and np.array_equal(vs[k], v1)
        assert np.array_equal(r["epot"][:, k], r1["epot"])
    # the stopped system stays still in the next call, which reports it again
    x_before = xs[1].copy()
    r = b.run(2, fixed_dt=dt, allow_stopped=True)
    assert np.all(r["dt"][:, 1] == 0.0)
    assert np.array_equal(_split(b.state()[0], [s.n for s in systems])[1], x_before)
    # without allow_stopped the same condition is an error naming the system
    with pytest.raises(EngineError, match=r"system 1"):
        b.run(1, fixed_dt=dt)
    # a single-system force call fails outright
    h = Engine.from_system(_hedgehog(), potentials=CHAIN, etors_compat=False)
    with pytest.raises(EngineError, match="16 bonds"):
        h.energy_grad()


def _ar_near_cluster(vz):
    cl = S.hydrocarbon_cluster(60, 5)
    cl.vel = S.maxwell_velocities(cl, 300.0, stream=6)
    top = cl.pos[np.argmax(cl.pos[:, 2])]
    ar = top + np.array([0.0, 0.0, 3.6])
    el = np.concatenate([cl.elem, [S.AR]]).astype(np.int32)
    pos = np.vstack([cl.pos, ar])
    vel = np.vstack([cl.vel, [[0.0, 0.0, vz]]])
    return S.System(el, pos, cl.pbc.copy(), vel, name="cluster+Ar")


@pytest.mark.parametrize("chain", [CHAIN, ("FBZL",) + CHAIN])
def test_dt_rule_has_neighbours_by_member(chain):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle
    vz = 1.0  # A/fs upwards, away from the cluster: 0.05 A / vz = 5e-17 s
    s = _ar_near_cluster(vz)
    mask = sum({"FBZL": 1, "REBO": 4, "AIREBO": 8, "ETors": 16}[p] for p in chain)
    o = Oracle.from_system(s, pots=mask, etors_compat=False)
    g = Engine.from_system(s, potentials=chain, etors_compat=False)
    ro, rg = o.run(6), g.run(6)
    print(chain, "dt oracle", ro["dt"], "gpu", rg["dt"], flush=True)
    assert np.allclose(rg["dt"], ro["dt"], rtol=1e-12, atol=0)
    # the dt after the last step, from the velocities the step left
    _, v, _, _ = g.state()
    speed = np.linalg.norm(v, axis=1)  # cm/s
    with_ar = "FBZL" in chain
    vmax = speed.max() if with_ar else speed[:-1].max()
    want = min(5e-16, 0.05 * ANG / vmax)
    assert g.time()[0, 2] == pytest.approx(want, rel=1e-12)
    assert (want < 1e-16) == with_ar  # the Ar, and only it, sets dt this low
