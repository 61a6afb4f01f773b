"""REBO bonds with LDS-staged shells (k_rebo_shell, tuning value occ_bonds
11/12): one wavefront per tile of 64 bonds stages the shells of the tile's
atoms in LDS.  Against the oracle on the parity inputs (FP64 bar: energies
1e-12, forces 1e-10 of the largest), on a jittered diamond whose tiles
overflow the staging capacity (those atoms are read from global memory), on
a bombardment frame at the ion's closest approach, over a 30-step
trajectory, and in the mixed form (force RMS <= 1e-5 eV/A of the oracle)
(GPU)."""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, EV_PER_A, force_rel_err, systems_for_parity

pytestmark = pytest.mark.gpu

NAMES = {"FBZL": 1, "REBO": 4, "AIREBO": 8, "ETors": 16}


def _engine(sys_, pots, precision, occ):
    from mdtk_amd import Engine
    e = Engine.from_system(sys_, potentials=pots, precision=precision)
    e.set_tune(dict(occ_bonds=occ, occ_bstar=1, lj_shape=3))
    return e


def _oracle(sys_, pots):
    from oracle.oracle import Oracle
    mask = 0
    for p in pots:
        mask |= NAMES[p]
    return Oracle.from_system(sys_, pots=mask)


def _cases():
    out = {k: v for k, v in systems_for_parity().items() if "REBO" in v[1]["pots"]}
    out["diamond512_jit"] = (S.jitter(S.diamond(4, 4, 4), 0.05, 21), dict(pots=("REBO",)))
    t = S.bombard_target(600, 3, height=2.0)
    t.pos[-1, 2] -= 3.2  # the ion inside the surface layer
    out["impact"] = (t, dict(pots=("FBZL", "REBO", "AIREBO", "ETors")))
    return out


@pytest.mark.parametrize("name", list(_cases().keys()))
def test_shell_fp64_matches_oracle(name):
    sys_, kw = _cases()[name]
    eo, go = _oracle(sys_, kw["pots"]).energy_grad()
    eg, gg = _engine(sys_, kw["pots"], "fp64", 11).energy_grad()
    print(f"{name} n={sys_.n}: E_oracle(eV)={eo / EV} E_shell(eV)={eg / EV} force rel err "
          f"{force_rel_err(gg, go):.2e}", flush=True)
    for a, b in zip(eo, eg):
        assert abs(a - b) <= 1e-12 * abs(a) + 1e-9 * EV
    assert force_rel_err(gg, go) <= 1e-10


@pytest.mark.parametrize("occ", [11, 12])
@pytest.mark.parametrize("name", ["pe360", "graphite_jit", "impact", "diamond512_jit"])
def test_shell_mixed_matches_oracle(name, occ):
    sys_, kw = _cases()[name]
    eo, go = _oracle(sys_, kw["pots"]).energy_grad()
    eg, gg = _engine(sys_, kw["pots"], "mixed", occ).energy_grad()
    _, g1 = _engine(sys_, kw["pots"], "mixed", 1).energy_grad()
    rms = np.sqrt((np.linalg.norm(gg - go, axis=1) ** 2).mean()) / EV_PER_A
    rms1 = np.sqrt((np.linalg.norm(gg - g1, axis=1) ** 2).mean()) / EV_PER_A
    print(f"{name} occ {occ}: mixed shell force RMS vs oracle {rms:.2e} eV/A, vs the per-bond kernel {rms1:.2e}",
          flush=True)
    assert rms <= 1e-5
    assert rms1 <= 1e-6


def test_shell_trajectory_fp64():
    from oracle.oracle import Oracle, POT_AIREBO, POT_ETORS, POT_REBO
    s = S.jitter(S.polyethylene(3, 3, 6), 0.04, 5)
    s.vel = S.maxwell_velocities(s, 600.0, stream=9)
    e = _engine(s, ("REBO", "AIREBO", "ETors"), "fp64", 11)
    o = Oracle.from_system(s, pots=POT_REBO | POT_AIREBO | POT_ETORS)
    rg = e.run(30, fixed_dt=0.2e-15)
    ro = o.run(30, fixed_dt=0.2e-15)
    d = np.abs(rg["epot"] - ro["epot"]).max() / np.abs(ro["epot"]).max()
    print(f"30 steps with LDS shells: max |dEpot|/|Epot| = {d:.2e}", flush=True)
    assert d <= 1e-10
