"""Differentiation form on the GPU: parameter overrides and dE/dtheta.

* Energies and forces at perturbed parameters equal the oracle's at the same
  parameters (1e-12 relative energy, 1e-10 relative force): pins the tables'
  rebuild on the device.
* dE_s/dtheta from the adjoint kernels equals the central finite difference
  of the GPU energy, h = 1e-6 relative (absolute for parameters that are 0):
  |adjoint - FD| <= 1e-6 max|dE/dtheta| per system.
* The torch.autograd wrappers: E.sum().backward() and the force-matching loss
  gradient against central differences of the loss over theta, and the
  loss's mixed derivative against the closed form of an LJ-only system.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, ANG, force_rel_err

pytestmark = pytest.mark.gpu

POTS = ("REBO", "AIREBO")


def _perturb(theta, seed, scale=0.01):
    u = S.uniforms(theta.size, 7000 + seed)
    return theta * (1 + scale * (2 * u - 1))


def _clusters(k, seed=0, lo=64, hi=200):
    u = S.uniforms(k, 8100 + seed)
    return [S.hydrocarbon_cluster(int(lo + (hi - lo) * u[i]), 300 + seed * 50 + i) for i in range(k)]


def test_perturbed_params_match_oracle():
    from mdtk_amd import Engine
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO
    for sys in (_clusters(1, 1)[0], S.jitter(S.polyethylene(2, 3, 5), 0.05, 4)):
        g = Engine.from_system(sys, potentials=POTS)
        o = Oracle.from_system(sys, pots=POT_REBO | POT_AIREBO)
        assert np.array_equal(o.params(), g.params())
        th = _perturb(g.params(), 1)
        g.set_params(th)
        o.set_params(th)
        eg, gg = g.energy_grad()
        eo, go = o.energy_grad()
        print(f"{sys.name}: E_gpu {eg / EV} E_oracle {eo / EV}", flush=True)
        for a, b in zip(eg, eo):
            assert abs(a - b) <= 1e-12 * abs(b) + 1e-9 * EV
        assert force_rel_err(gg, go) <= 1e-10


def _fd_step(v):
    return 1e-6 * abs(v) if v != 0 else None


def test_dE_dtheta_matches_central_differences():
    from mdtk_amd import Engine
    from mdtk_amd.engine import PARAM_NAMES
    mem = _clusters(3, 2)
    eng = Engine.from_systems(mem, potentials=POTS)
    th0 = eng.params()
    e, g, d = eng.energy_grad_params()
    abs_steps = {"B": 1e-6 * EV, "beta": 1e-6 / ANG}
    steps = np.array([_fd_step(v) or abs_steps["beta" if n.startswith("beta") else "B"]
                      for v, n in zip(th0, PARAM_NAMES)])
    # sensitivity to a relative change of each parameter (1e6 steps for zeros)
    sens = np.abs(d) * (steps * 1e6)
    scale = sens.max(1)
    for k, name in enumerate(PARAM_NAMES):
        h = steps[k]
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        eng.set_params(tp)
        ep = eng.energy_grad()[0].sum(1)
        eng.set_params(tm)
        em = eng.energy_grad()[0].sum(1)
        fd = (ep - em) / (2 * h)
        err = np.abs(fd - d[:, k]) * (h * 1e6) / scale
        print(f"{name}: adjoint {d[:, k]} fd {fd} rel err {err.max():.2e}", flush=True)
        assert err.max() <= 1e-6, name
    eng.set_params(th0)


def test_energy_function_backward():
    import torch
    from mdtk_amd import Engine
    from mdtk_amd.autograd import energy
    mem = _clusters(2, 3)
    eng = Engine.from_systems(mem, potentials=POTS)
    x0 = np.concatenate([s.to_cgs()[1].reshape(-1, 3) for s in mem])
    th0 = eng.params()
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    w = torch.tensor([0.7, -1.3], dtype=torch.float64)
    E = energy(eng, x, th)
    (E * w).sum().backward()
    _, g, d = eng.energy_grad_params()
    off = eng.offsets
    want_x = np.concatenate([g[off[s]:off[s + 1]] * w[s].item() for s in range(2)])
    # a second force call: equal up to the order of the FP64 atomic sums
    assert np.abs(x.grad.numpy() - want_x).max() <= 1e-12 * np.abs(want_x).max()
    want_t = (d * w.numpy()[:, None]).sum(0)
    assert np.all(np.abs(th.grad.numpy() - want_t) <= 1e-12 * np.abs(d).max(0) + 1e-300)


def test_force_matching_loss_gradient():
    """Loss = sum (E - E_ref)^2 + sum |F - F_ref|^2 against targets made at
    perturbed parameters; dLoss/dtheta from the op against central
    differences of the loss (h = 1e-6 relative)."""
    import torch
    from mdtk_amd import Engine
    from mdtk_amd.autograd import force_matching_loss
    mem = _clusters(2, 4)
    eng = Engine.from_systems(mem, potentials=POTS)
    x0 = np.concatenate([s.to_cgs()[1].reshape(-1, 3) for s in mem])
    th0 = eng.params()
    eng.set_params(_perturb(th0, 5))
    eng.set_positions(x0)
    e_ref, g_ref = eng.energy_grad()
    e_ref = torch.tensor(e_ref.sum(1))
    f_ref = torch.tensor(-g_ref)
    x = torch.tensor(x0)
    th = torch.tensor(th0, requires_grad=True)
    L = force_matching_loss(eng, x, th, e_ref, f_ref)
    L.backward()
    gth = th.grad.numpy()
    checked = 0
    for k in range(th0.size):
        if th0[k] == 0:
            continue
        h = 1e-6 * abs(th0[k])
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        lp = force_matching_loss(eng, x, torch.tensor(tp), e_ref, f_ref).item()
        lm = force_matching_loss(eng, x, torch.tensor(tm), e_ref, f_ref).item()
        fd = (lp - lm) / (2 * h)
        # compare sensitivities to a relative change of each parameter
        rel = abs(fd - gth[k]) * abs(th0[k]) / np.abs(gth * th0).max()
        print(f"theta[{k}]: op {gth[k]:.6e} fd {fd:.6e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-5
        checked += 1
    assert checked >= 20


def test_force_matching_mixed_derivative_closed_form():
    """The mixed derivative d(w . F)/dtheta against a closed form: carbon
    atoms 3.9-5.3 A apart (no REBO bonds, S_r = 0, C_ij = 1) interact by
    E = sum f(r) 4 eps ((s/r)^12 - (s/r)^6) alone, so F is proportional to
    eps_CC and dF/deps_CC = F / eps_CC exactly.  With E_ref = E and F_ref = 0
    the loss is L = sum |F|^2 and dL/deps_CC = 2 L / eps_CC; the op's value
    (five-point difference of the exact adjoint) agrees to 1e-9."""
    import torch
    from mdtk_amd import Engine
    from mdtk_amd.engine import PARAM_NAMES
    from mdtk_amd.autograd import force_matching_loss
    pos = np.array([[0.0, 0.0, 0.0], [3.9, 0.0, 0.0], [3.9, 4.2, 0.3], [0.2, 4.6, 2.9], [8.1, 1.0, 0.7]])
    sys_ = S.System(np.full(len(pos), S.C, np.int32), pos, np.full(3, S.NO_PBC), None, None, "c5")
    eng = Engine.from_system(sys_, potentials=POTS)
    k = PARAM_NAMES.index("eps_CC")
    th0 = eng.params()
    x0 = sys_.to_cgs()[1].reshape(-1, 3)
    e0, _ = eng.energy_grad()
    e_ref = torch.tensor([np.sum(e0)])  # one system
    f_ref = torch.zeros((len(pos), 3), dtype=torch.float64)
    th = torch.tensor(th0, requires_grad=True)
    L = force_matching_loss(eng, torch.tensor(x0), th, e_ref, f_ref)
    L.backward()
    want = 2.0 * L.item() / th0[k]
    got = th.grad.numpy()[k]
    print(f"  L = {L.item():.6e} (eV/A)^2, dL/deps_CC op {got:.12e} closed form {want:.12e}", flush=True)
    assert L.item() > 1e-6
    assert abs(got - want) <= 1e-9 * abs(want)


def test_c4_scale_dE_dtheta():
    """The differentiation config at its own size: 512 seeded hydrocarbon
    clusters of 64-512 atoms in one batch.  The adjoint dE_s/dtheta of every
    system equals the central difference of the batched GPU energy (h = 1e-6
    relative, as above) for every parameter, and two members' energies and
    forces equal the oracle's."""
    from mdtk_amd import Engine
    from mdtk_amd.engine import PARAM_NAMES
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO
    mem = _clusters(512, 9, lo=64, hi=512)
    eng = Engine.from_systems(mem, potentials=POTS)
    th0 = eng.params()
    e, g, d = eng.energy_grad_params()
    off = eng.offsets
    for k in (0, 511):
        eo, go = Oracle.from_system(mem[k], pots=POT_REBO | POT_AIREBO).energy_grad()
        assert np.abs(e[k] - eo).max() <= 1e-12 * np.abs(eo).max()
        assert force_rel_err(g[off[k]:off[k + 1]], go) <= 1e-10
    abs_steps = {"B": 1e-6 * EV, "beta": 1e-6 / ANG}
    steps = np.array([_fd_step(v) or abs_steps["beta" if n.startswith("beta") else "B"]
                      for v, n in zip(th0, PARAM_NAMES)])
    sens = np.abs(d) * (steps * 1e6)
    scale = sens.max(1)
    worst = 0.0
    for k, name in enumerate(PARAM_NAMES):
        h = steps[k]
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        eng.set_params(tp)
        ep = eng.energy_grad()[0].sum(1)
        eng.set_params(tm)
        em = eng.energy_grad()[0].sum(1)
        err = np.abs((ep - em) / (2 * h) - d[:, k]) * (h * 1e6) / scale
        worst = max(worst, err.max())
        assert err.max() <= 1e-6, name
    eng.set_params(th0)
    print(f"512 clusters, {sum(s.n for s in mem)} atoms: worst adjoint-FD error {worst:.2e}", flush=True)
