"""Batched form on the GPU: many independent systems in one context.

Each system of a batch must evolve exactly as it does alone: energies per
chain member, forces, neighbour lists, the skin-test rebuild schedule and the
adaptive dt are per system.  Batch against single-system contexts (same
kernels, one system each): energies |dE| <= 1e-12 |E|, forces max|dF| /
max|F| <= 1e-12 (only the order of the FP64 atomic sums differs), and over a
trajectory the per-step energies within 1e-10 relative.  One member is also
checked against the oracle directly.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, force_rel_err, pe_small

pytestmark = pytest.mark.gpu

POTS = ("REBO", "AIREBO", "ETors")


def _members():
    a = pe_small(seed=11, jit=0.05)
    a.vel = S.maxwell_velocities(a, 300.0, stream=21)
    b = S.hydrocarbon_cluster(90, 5)
    b.vel = S.maxwell_velocities(b, 600.0, stream=22)
    c = S.jitter(S.graphite(5, 3, 2), 0.04, 8)
    c.vel = S.maxwell_velocities(c, 300.0, stream=23)
    d = S.bombard_target(700, 24)
    return [a, b, c, d]


def _split(g, off):
    return [g[off[i]:off[i + 1]] for i in range(len(off) - 1)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_energy_grad_matches_single(precision):
    from mdtk_amd import Engine
    pots = ("FBZL",) + POTS
    mem = _members()
    bat = Engine.from_systems(mem, potentials=pots, precision=precision)
    eb, gb = bat.energy_grad()
    assert eb.shape == (len(mem), len(pots))
    for k, (s, g) in enumerate(zip(mem, _split(gb, bat.offsets))):
        one = Engine.from_system(s, potentials=pots, precision=precision)
        e1, g1 = one.energy_grad()
        print(f"{precision} member {k} ({s.name}, n={s.n}): E={eb[k].sum() / EV:.9f} eV", flush=True)
        for a, b in zip(eb[k], e1):
            assert abs(a - b) <= 1e-12 * abs(b) + 1e-12 * EV
        assert force_rel_err(g, g1) <= 1e-12


def test_batch_member_matches_oracle():
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO, POT_ETORS
    mem = _members()
    bat = Engine.from_systems(mem, potentials=POTS)
    eb, gb = bat.energy_grad()
    k = 1
    o = Oracle.from_system(mem[k], pots=POT_REBO | POT_AIREBO | POT_ETORS)
    eo, go = o.energy_grad()
    for a, b in zip(eb[k], eo):
        assert abs(a - b) <= 1e-12 * abs(b) + 1e-9 * EV
    assert force_rel_err(_split(gb, bat.offsets)[k], go) <= 1e-10


def test_batch_neighbour_lists():
    from mdtk_amd import Engine
    mem = _members()[:3]
    bat = Engine.from_systems(mem, potentials=POTS)
    off = bat.offsets
    for term in range(len(POTS)):
        lb = bat.neighbour_lists(term)
        for k, s in enumerate(mem):
            one = Engine.from_system(s, potentials=POTS)
            l1 = one.neighbour_lists(term)
            for i in range(s.n):
                assert np.array_equal(lb[off[k] + i] - off[k], l1[i])


@pytest.mark.parametrize("fixed_dt", [0.0, 0.2e-15])
def test_batch_trajectory_matches_single(fixed_dt):
    """40 steps: per-system energies, dt and final state equal the single runs;
    with the adaptive rule every system keeps its own dt."""
    from mdtk_amd import Engine
    pots = ("FBZL",) + POTS
    mem = _members()
    bat = Engine.from_systems(mem, potentials=pots)
    rb = bat.run(40, fixed_dt=fixed_dt)
    xb = bat.state()[0]
    assert rb["epot"].shape == (40, len(mem))
    for k, s in enumerate(mem):
        one = Engine.from_system(s, potentials=pots)
        r1 = one.run(40, fixed_dt=fixed_dt)
        et_b = rb["epot"][:, k] + rb["ekin"][:, k]
        et_1 = r1["epot"] + r1["ekin"]
        print(f"dt={fixed_dt} member {k}: max|dE_tot| {np.abs(et_b - et_1).max() / EV:.3e} eV, "
              f"dt range {rb['dt'][:, k].min():.3e}..{rb['dt'][:, k].max():.3e}", flush=True)
        assert np.abs(et_b - et_1).max() <= 1e-10 * np.abs(et_1).max() + 1e-9 * EV
        assert np.allclose(rb["dt"][:, k], r1["dt"], rtol=1e-10, atol=0)
        x1 = one.state()[0]
        assert np.abs(xb[bat.offsets[k]:bat.offsets[k + 1]] - x1).max() <= 1e-9 * S.ANG
    if fixed_dt == 0.0:
        assert len({float(v) for v in rb["dt"][-1]}) > 1  # per-system adaptive dt


def test_batch_uneven_bombardment_runs():
    """A small uneven bombardment batch (with fixed atoms and the Ar ion)
    steps with the graph and conserves each member's total energy at 0.1 fs
    (exact torsion adjoint: the reference's torsion derivative is not
    conservative)."""
    from mdtk_amd import Engine
    mem = [S.bombard_target(n, t) for t, n in enumerate((600, 1500, 900))]
    bat = Engine.from_systems(mem, potentials=("FBZL",) + POTS, precision="mixed", etors_compat=False)
    r = bat.run(60, fixed_dt=0.1e-15)
    et = (r["epot"] + r["ekin"]) / EV
    for k, s in enumerate(mem):
        drift = np.abs(et[:, k] - et[0, k]).max() / s.n
        print(f"member {k} n={s.n}: max |dE|/atom {drift:.3e} eV", flush=True)
        assert drift < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_keeps_results(precision):
    """The autotuner changes launch shapes only: forces and a trajectory after
    tuning equal an untuned engine's (same kernels, other grid sizes and
    register budgets; only the order of FP64 atomic sums may differ)."""
    from mdtk_amd import Engine
    pots = ("FBZL",) + POTS
    mem = _members()
    a = Engine.from_systems(mem, potentials=pots, precision=precision)
    b = Engine.from_systems(mem, potentials=pots, precision=precision)
    rep = b.autotune(reps=2)
    print(f"{precision} autotune: {rep}", flush=True)
    ea, ga = a.energy_grad()
    eb, gb = b.energy_grad()
    assert np.abs(ea - eb).max() <= 1e-12 * np.abs(ea).max()
    assert force_rel_err(gb, ga) <= 1e-12
    ra = a.run(20, fixed_dt=0.2e-15)
    rb = b.run(20, fixed_dt=0.2e-15)
    assert np.abs(ra["epot"] - rb["epot"]).max() <= 1e-10 * np.abs(ra["epot"]).max()
