"""Mixed-precision form and graph replay, on the GPU.

Mixed precision (FP32 displacements, cutoffs and splines in Angstrom/eV; FP64
coordination and bond-order sums, energies and force accumulation) is checked
against the FP64 engine: per-atom force RMS error <= 1e-5 eV/A (the mixed
form's accuracy bar), energy within 1e-6 relative.
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
