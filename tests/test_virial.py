"""Virial W = sum over gradient pushes of -r (x) v (no reference counterpart:
MDTK computes none).

CPU pins of the oracle's virial, each independent of how it is summed:
* an open system has no images, so W must equal -sum_i x_i (x) dE/dx_i of
  the oracle's own gradient (every push is balanced; a push left out of the
  virial, or counted twice, breaks the identity);
* a periodic system: the diagonal components are minus the derivative of the
  energy under a homogeneous strain of the atoms and the box, W_aa =
  -dE/d(eps_aa), by central differences (exact adjoint forces; the torsion
  compat derivative and the b* r_alter derivative are not gradients);
* the fcc metals at their tabulated lattice constants sit at zero pressure
  (the Ackland fit, Ackland.cxx:304,327,350).
GPU: the engine's in-kernel virial equals the oracle's (FP64, 1e-10 of its
largest component) on open and periodic systems of every chain member, in
batches, per step inside mdg_step_nve, and in the mixed form within its
force-accuracy bound.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from oracle import oracle as O
from tests.helpers import ANG, EV, pe_small

POTS = O.POT_REBO | O.POT_AIREBO | O.POT_ETORS


def _identity(sys_, pots, **kw):
    o = O.Oracle.from_system(sys_, pots=pots, **kw)
    _, g = o.energy_grad()
    w = o.virial()
    x = sys_.pos * ANG
    full = -(x[:, :, None] * g[:, None, :]).sum(0)
    want = np.array([full[0, 0], full[1, 1], full[2, 2], full[0, 1], full[0, 2], full[1, 2]])
    return w, want


@pytest.mark.parametrize("case", ["hydrocarbon", "hydrocarbon_compat", "ethane_brenner", "ion_slab_open",
                                  "metals"])
def test_oracle_virial_open_system_identity(case):
    if case.startswith("hydrocarbon"):
        s, pots, kw = S.hydrocarbon_cluster(80, 4), POTS, dict(etors_compat=case.endswith("compat"))
    elif case == "ethane_brenner":
        s, pots, kw = S.jitter(S.ethane(), 0.05, 3), O.POT_BRENNER, {}
    elif case == "ion_slab_open":
        t = S.bombard_target(600, 2)
        t.pos[-1, 2] -= 7.0  # the ion inside the ZBL range of the surface
        s = S.System(t.elem, t.pos, np.array([S.NO_PBC] * 3), name="ionslab")
        pots, kw = O.POT_FBZL | POTS, {}
    else:
        rng = np.random.default_rng(5)
        p = rng.uniform(0, 9, (40, 3))
        keep = [0]
        for i in range(1, len(p)):
            if np.min(np.linalg.norm(p[keep] - p[i], axis=1)) > 2.0:
                keep.append(i)
        p = p[keep]
        el = rng.choice([64, 108, 197, 1, 12, 40], len(p)).astype(np.int32)
        s = S.System(el, p + 20.0, np.array([S.NO_PBC] * 3))
        pots, kw = O.POT_ACKLAND | O.POT_FLJ | O.POT_FBM | O.POT_TB | O.POT_FBZL, {}
    w, want = _identity(s, pots, **kw)
    assert np.abs(want).max() > 0
    assert np.abs(w - want).max() <= 1e-11 * np.abs(want).max(), (w, want)


def _strained_energy(sys_, pots, eps, axis):
    p = sys_.pos.copy()
    L = sys_.pbc.copy()
    p[:, axis] *= (1 + eps)
    L[axis] *= (1 + eps)
    e, _ = O.Oracle.from_system(S.System(sys_.elem, p, L), pots=pots, etors_compat=False).energy_grad()
    return e.sum()


@pytest.mark.parametrize("case", ["pe", "diamond_brenner", "cu"])
def test_oracle_virial_is_strain_derivative(case):
    if case == "pe":
        s, pots = S.jitter(S.polyethylene(2, 3, 5), 0.06, 9), POTS
    elif case == "diamond_brenner":
        s, pots = S.jitter(S.diamond(3, 3, 3), 0.05, 8), O.POT_BRENNER
    else:
        s, pots = S.jitter(S.fcc(64, 4, 4, 4), 0.05, 7), O.POT_ACKLAND
    O.set_alter_exact(True)
    try:
        w = O.Oracle.from_system(s, pots=pots, etors_compat=False).virial()
        h = 1e-6
        for a in range(3):
            fd = -(_strained_energy(s, pots, h, a) - _strained_energy(s, pots, -h, a)) / (2 * h)
            assert abs(w[a] - fd) <= 2e-7 * np.abs(w[:3]).max(), (case, a, w[a], fd)
    finally:
        O.set_alter_exact(False)


@pytest.mark.parametrize("el", [64, 108, 197])
def test_oracle_fcc_zero_pressure(el):
    s = S.fcc(el, 4, 4, 4)
    w = O.Oracle.from_system(s, pots=O.POT_ACKLAND).virial()
    vol = np.prod(s.pbc * ANG)
    p_gpa = w[:3].sum() / (3 * vol) * 1e-10
    assert abs(p_gpa) < 2e-3  # the bulk moduli are 100-180 GPa
    assert np.abs(w[3:]).max() <= 1e-10 * np.abs(w[:3]).max() + 1e-30  # cubic: no shear


# ------------------------------------------------------------------- GPU
def _gpu_vs_oracle(sys_, names, pots, precision="fp64", tol=1e-10, **kw):
    from mdtk_amd import Engine
    e = Engine.from_system(sys_, potentials=names, precision=precision, **kw)
    e.set_virial(True)
    e.energy_grad()
    wg = e.virial()
    wo = O.Oracle.from_system(sys_, pots=pots, **kw).virial()
    err = np.abs(wg - wo).max() / np.abs(wo).max()
    print(f"virial {sys_.name} {precision}: rel err {err:.2e}", flush=True)
    assert err <= tol, (wg, wo)


GPU_CASES = ["pe", "cluster", "graphite", "bombard", "brenner", "metals"]


def _case(name):
    if name == "pe":
        return S.jitter(S.polyethylene(3, 4, 6), 0.07, 3), ("REBO", "AIREBO", "ETors"), POTS
    if name == "cluster":
        return S.hydrocarbon_cluster(120, 5), ("REBO", "AIREBO", "ETors"), POTS
    if name == "graphite":
        return S.jitter(S.graphite(5, 3, 2), 0.05, 4), ("REBO", "AIREBO", "ETors"), POTS
    if name == "bombard":
        t = S.bombard_target(1200, 6)
        t.pos[-1, 2] -= 7.5
        return t, ("FBZL", "REBO", "AIREBO", "ETors"), O.POT_FBZL | POTS
    if name == "brenner":
        return S.jitter(S.diamond(3, 3, 3), 0.06, 2), ("Brenner",), O.POT_BRENNER
    rng = np.random.default_rng(11)
    cu = S.jitter(S.fcc(64, 5, 5, 5), 0.06, 9)
    el = cu.elem.copy()
    el[rng.choice(len(el), 40, replace=False)] = 197
    return S.System(el, cu.pos, cu.pbc, name="cuau"), ("Ackland", "FBM", "TightBinding"), \
        O.POT_ACKLAND | O.POT_FBM | O.POT_TB


@pytest.mark.gpu
@pytest.mark.parametrize("name", GPU_CASES)
def test_gpu_virial_matches_oracle(name):
    s, names, pots = _case(name)
    _gpu_vs_oracle(s, names, pots, etors_compat=False)


@pytest.mark.gpu
def test_gpu_virial_mixed_within_force_accuracy():
    s, names, pots = _case("pe")
    _gpu_vs_oracle(s, names, pots, precision="mixed", tol=1e-6, etors_compat=False)


@pytest.mark.gpu
def test_gpu_virial_batch_and_steps():
    """A batch returns each system's virial (equal to the system alone), and
    mdg_step_nve reports the virial of every step's force call."""
    from mdtk_amd import Engine
    systems = [_case("pe")[0], _case("bombard")[0], _case("cluster")[0]]
    names = ("FBZL", "REBO", "AIREBO", "ETors")
    b = Engine.from_systems(systems, potentials=names, etors_compat=False)
    b.set_virial(True)
    b.energy_grad()
    wb = b.virial()
    for k, s in enumerate(systems):
        e = Engine.from_system(s, potentials=names, etors_compat=False)
        e.set_virial(True)
        e.energy_grad()
        w1 = e.virial()
        assert np.abs(wb[k] - w1).max() <= 1e-12 * np.abs(w1).max()
    r = b.run(3, fixed_dt=0.1e-15)
    assert r["virial"].shape == (3, 3, 6)
    b.energy_grad()  # at the positions after the third step
    assert np.abs(r["virial"][-1] - b.virial()).max() <= 1e-12 * np.abs(b.virial()).max()


@pytest.mark.gpu
def test_gpu_virial_fresh_engines_first_call():
    """The first force call of each new context (lists built, cold caches)
    gives the oracle's virial, and the next call repeats its gradient bit for
    bit (the FP64 form is deterministic)."""
    from mdtk_amd import Engine
    s, _, _ = _case("pe")
    wo = O.Oracle.from_system(s, pots=O.POT_REBO, etors_compat=False).virial()
    for _ in range(6):
        e = Engine.from_system(s, potentials=("REBO",), etors_compat=False)
        e.set_virial(True)
        _, g1 = e.energy_grad()
        w1 = e.virial()
        _, g2 = e.energy_grad()
        assert np.abs(w1 - wo).max() <= 1e-12 * np.abs(wo).max()
        assert np.array_equal(g1, g2)
