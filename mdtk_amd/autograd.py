"""The energy-with-backward op as torch.autograd Functions.

``energy(engine, x, theta)`` returns the per-system potential energies E_s
(erg) of an FP64 engine at positions x (n x 3, cm; the systems' atoms one after
another) and parameters theta (NUM_PARAMS, CGS; layout in
``engine.PARAM_NAMES``).  Its backward is the engine's adjoint kernels:
dE/dx (the force call's gradient) and dE_s/dtheta (the parameter adjoint of
the REBO pair and LJ terms).  Every evaluation pushes the positions with a
forced neighbour-list rebuild, so neighbour shells are recomputed, never
stored between calls.

``energy_forces(engine, x, theta)`` returns (E, F = -dE/dx) and is
differentiable to second order in the way a force-matching loss needs: the
backward of F contracts the mixed derivative d^2E/dx dtheta (and the Hessian
d^2E/dx^2) with the incoming force cotangent w.  That contraction is the
directional derivative of the exact first-order adjoint along w, taken by
the five-point central difference over the positions, x +- h w and
x +- 2h w with max|h w| = 1e-4 Angstrom (truncation error O(h^4) relative to
the adjoint itself).
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import NUM_PARAMS

_ANG = 1e-8
_EV = 1.6021e-12
FD_STEP_CM = 1e-4 * _ANG


def _evaluate(engine, x: np.ndarray, theta: np.ndarray):
    engine.set_params(theta)
    engine.set_positions(x, dr_cm=None)
    e, g, d = engine.energy_grad_params()
    m = max(engine.nsys, 1)
    e = np.asarray(e).reshape(m, -1).sum(1)
    d = np.asarray(d).reshape(m, NUM_PARAMS)
    return e, g, d


def _rows_scale(engine, g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Scales the rows of system s by w[s]."""
    off = engine.offsets
    counts = np.diff(off)
    return g * np.repeat(w, counts)[:, None]


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, theta, engine):
        xn = x.detach().cpu().double().numpy().reshape(-1, 3)
        tn = theta.detach().cpu().double().numpy()
        e, g, d = _evaluate(engine, xn, tn)
        ctx.engine = engine
        ctx.save_for_backward(torch.from_numpy(g), torch.from_numpy(d))
        ctx.dev = x.device
        return torch.from_numpy(e).to(x.device)

    @staticmethod
    def backward(ctx, gE):
        g, d = ctx.saved_tensors
        w = gE.detach().cpu().double().numpy()
        gx = _rows_scale(ctx.engine, g.numpy(), w)
        gth = (d.numpy() * w[:, None]).sum(0)
        return (torch.from_numpy(gx).to(ctx.dev), torch.from_numpy(gth).to(ctx.dev), None)


def energy(engine, x: torch.Tensor, theta: torch.Tensor) -> torch.Tensor:
    """Per-system energies E_s (erg), differentiable in x and theta."""
    return _Energy.apply(x, theta, engine)


class _EnergyForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, theta, engine):
        xn = x.detach().cpu().double().numpy().reshape(-1, 3)
        tn = theta.detach().cpu().double().numpy()
        e, g, d = _evaluate(engine, xn, tn)
        ctx.engine = engine
        ctx.xn, ctx.tn = xn, tn
        ctx.save_for_backward(torch.from_numpy(g), torch.from_numpy(d))
        ctx.dev = x.device
        return torch.from_numpy(e).to(x.device), torch.from_numpy(-g).to(x.device)

    @staticmethod
    def backward(ctx, gE, gF):
        g, d = ctx.saved_tensors
        eng = ctx.engine
        m = max(eng.nsys, 1)
        wE = np.zeros(m) if gE is None else gE.detach().cpu().double().numpy()
        gx = _rows_scale(eng, g.numpy(), wE)
        gth = (d.numpy() * wE[:, None]).sum(0)
        if gF is not None:
            w = gF.detach().cpu().double().numpy().reshape(-1, 3)
            wmax = np.abs(w).max()
            if wmax > 0:
                h = FD_STEP_CM / wmax
                _, gp, dp = _evaluate(eng, ctx.xn + h * w, ctx.tn)
                _, gm, dm = _evaluate(eng, ctx.xn - h * w, ctx.tn)
                _, gp2, dp2 = _evaluate(eng, ctx.xn + 2 * h * w, ctx.tn)
                _, gm2, dm2 = _evaluate(eng, ctx.xn - 2 * h * w, ctx.tn)
                # F = -dE/dx: d(w . F)/dtheta = -(w . d/dx) dE/dtheta;
                # f'(0) = (8 (f(h) - f(-h)) - (f(2h) - f(-2h))) / 12h
                gth -= (8 * (dp - dm) - (dp2 - dm2)).sum(0) / (12 * h)
                gx -= (8 * (gp - gm) - (gp2 - gm2)) / (12 * h)
                eng.set_positions(ctx.xn, dr_cm=None)
        return (torch.from_numpy(gx).to(ctx.dev), torch.from_numpy(gth).to(ctx.dev), None)


def energy_forces(engine, x: torch.Tensor, theta: torch.Tensor):
    """(E_s per system, F = -dE/dx), differentiable in theta (and x) for
    force-matching losses."""
    return _EnergyForces.apply(x, theta, engine)


def force_matching_loss(engine, x, theta, e_ref, f_ref):
    """sum_s (E_s - E_ref,s)^2 + sum_a |F_a - F_ref,a|^2 in eV^2 and (eV/A)^2;
    e_ref (erg) and f_ref (erg/cm) in the engine's CGS units."""
    e, f = energy_forces(engine, x, theta)
    return (((e - e_ref) / _EV) ** 2).sum() + (((f - f_ref) / (_EV / _ANG)) ** 2).sum()
