"""Mixed-precision form and graph replay, on the GPU.

Mixed precision (FP32 displacements, cutoffs and splines in Angstrom/eV; FP64
coordination and bond-order sums, energies and force accumulation) is judged
against the FP64 oracle directly (per-atom force RMS error <= 1e-5 eV/A, the
mixed form's accuracy bar, energy within 1e-6 relative) on polyethylene,
graphite and a bombardment frame at the ion's closest approach, and against
the FP64 engine on the 1,048,576-atom graphite system, where the oracle is
too slow; there, size-independent properties pin the FP64 engine: the
energy per atom of the perfect crystal equals that of a small cell (which
the oracle evaluates), and the net force of the periodic system is zero.
"""
import os

import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, EV_PER_A, pe_small

pytestmark = pytest.mark.gpu


def _engines(sys, pots):
    from mdtk_amd import Engine
    a = Engine.from_system(sys, potentials=pots)
    b = Engine.from_system(sys, potentials=pots, precision="mixed")
    return a, b


@pytest.mark.parametrize("name", ["pe", "cluster", "graphite"])
def test_mixed_force_rms(name):
    sys = {"pe": pe_small(), "cluster": S.hydrocarbon_cluster(60, 3),
           "graphite": S.jitter(S.graphite(5, 3, 2), 0.04, 5)}[name]
    a, b = _engines(sys, ("REBO", "AIREBO", "ETors"))
    ea, ga = a.energy_grad()
    eb, gb = b.energy_grad()
    rms = np.sqrt(((ga - gb) ** 2).sum(1).mean()) / EV_PER_A
    print(f"{name}: force RMS err {rms:.3e} eV/A, E64={ea.sum() / EV:.9f} Emix={eb.sum() / EV:.9f}", flush=True)
    assert rms <= 1e-5
    assert abs(ea.sum() - eb.sum()) <= 1e-6 * abs(ea.sum())


def _closest_approach():
    """A bombardment target with the 100 eV Ar 0.75 A above a surface carbon
    (about where the Ar-C ZBL energy reaches the ion's energy)."""
    t = S.bombard_target(600, 2)
    top = np.nonzero(t.elem[:-1] == S.C)[0]
    k = top[np.argmax(t.pos[top, 2])]
    t.pos[-1] = t.pos[k] + np.array([0.0, 0.0, 0.75])
    return t


@pytest.mark.parametrize("name", ["pe", "graphite", "collision"])
def test_mixed_forces_vs_oracle(name):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_FBZL, POT_REBO, POT_AIREBO, POT_ETORS
    pots, mask = ("REBO", "AIREBO", "ETors"), POT_REBO | POT_AIREBO | POT_ETORS
    if name == "pe":
        sys_ = S.jitter(S.polyethylene(3, 4, 6), 0.07, 3)
    elif name == "graphite":
        sys_ = S.jitter(S.graphite(5, 3, 2), 0.05, 4)
    else:
        sys_ = _closest_approach()
        pots, mask = ("FBZL",) + pots, mask | POT_FBZL
    eo, go = Oracle.from_system(sys_, pots=mask, etors_compat=False).energy_grad()
    em, gm = Engine.from_system(sys_, potentials=pots, precision="mixed", etors_compat=False).energy_grad()
    rms = np.sqrt(((gm - go) ** 2).sum(1).mean()) / EV_PER_A
    fmax = np.linalg.norm(go, axis=1).max() / EV_PER_A
    print(f"{name}: mixed vs oracle force RMS {rms:.3e} eV/A (max |F| {fmax:.1f} eV/A), "
          f"E {eo.sum() / EV:.9f} / {em.sum() / EV:.9f} eV", flush=True)
    assert rms <= 1e-5
    assert abs(em.sum() - eo.sum()) <= 1e-6 * abs(eo.sum())


def test_graphite_1m_properties():
    """The 1,048,576-atom AB graphite (64 x 64 x 32 cells): FP64 energy per
    atom of the perfect crystal equals the small cell's, which equals the
    oracle's; on the jittered crystal the net force is zero (periodic, no
    external field) and the mixed form's force RMS error against FP64 is
    within the bar."""
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO, POT_ETORS
    pots = ("REBO", "AIREBO", "ETors")
    small = S.graphite(5, 3, 2)
    eo, _ = Oracle.from_system(small, pots=POT_REBO | POT_AIREBO | POT_ETORS, etors_compat=False).energy_grad()
    es, _ = Engine.from_system(small, potentials=pots, etors_compat=False).energy_grad()
    assert abs(es.sum() - eo.sum()) <= 1e-12 * abs(eo.sum())
    big = S.graphite(64, 64, 32)
    assert big.n == 1048576
    eb, gb = Engine.from_system(big, potentials=pots, etors_compat=False).energy_grad()
    per_small, per_big = eo.sum() / small.n, eb.sum() / big.n
    print(f"E/atom: 1M {per_big / EV:.12f} eV, small cell {per_small / EV:.12f} eV", flush=True)
    assert abs(per_big - per_small) <= 1e-11 * abs(per_small)
    jit = S.jitter(big, 0.03, 9)
    e64, g64 = Engine.from_system(jit, potentials=pots, etors_compat=False).energy_grad()
    net = np.abs(g64.sum(0)).max() / EV_PER_A
    fmax = np.abs(g64).max() / EV_PER_A
    emx, gmx = Engine.from_system(jit, potentials=pots, etors_compat=False, precision="mixed").energy_grad()
    rms = np.sqrt(((gmx - g64) ** 2).sum(1).mean()) / EV_PER_A
    print(f"1M jittered: |sum F| {net:.3e} eV/A (max |F_i| {fmax:.2f}), mixed-FP64 RMS {rms:.3e} eV/A", flush=True)
    assert net <= 1e-9 * fmax * np.sqrt(big.n)
    assert rms <= 1e-5
    assert abs(emx.sum() - e64.sum()) <= 1e-6 * abs(e64.sum())


def test_mixed_nve_drift():
    """Exact torsion adjoint: the mixed engine's E(t) stays within a few meV/atom
    of the FP64 engine over 200 steps at 0.2 fs."""
    from mdtk_amd import Engine
    sys = pe_small(seed=3, jit=0.02)
    sys.vel = S.maxwell_velocities(sys, 300.0, stream=6)
    out = {}
    for prec in ("fp64", "mixed"):
        e = Engine.from_system(sys, potentials=("REBO", "AIREBO", "ETors"), etors_compat=False, precision=prec)
        r = e.run(200, fixed_dt=0.2e-15)
        out[prec] = (r["epot"] + r["ekin"]) / EV / sys.n
    d64 = out["fp64"] - out["fp64"][0]
    dmx = out["mixed"] - out["mixed"][0]
    print(f"max |dE| per atom: fp64 {np.abs(d64).max():.3e} eV, mixed {np.abs(dmx).max():.3e} eV", flush=True)
    assert np.abs(dmx - d64).max() < 1e-4


def test_graph_replay_matches_eager(monkeypatch):
    """The captured one-step hipGraph replays the same trajectory as eager launches."""
    from mdtk_amd import Engine
    sys = pe_small(seed=5, jit=0.02)
    sys.vel = S.maxwell_velocities(sys, 300.0, stream=7)
    g = Engine.from_system(sys, potentials=("REBO", "AIREBO", "ETors"))
    rg = g.run(30, fixed_dt=0.2e-15)
    monkeypatch.setenv("MDG_NO_GRAPH", "1")
    e = Engine.from_system(sys, potentials=("REBO", "AIREBO", "ETors"))
    re_ = e.run(30, fixed_dt=0.2e-15)
    assert np.abs(rg["epot"] - re_["epot"]).max() <= 1e-11 * np.abs(re_["epot"]).max()
    xg = g.state()[0]
    xe = e.state()[0]
    assert np.abs(xg - xe).max() <= 1e-9 * S.ANG
