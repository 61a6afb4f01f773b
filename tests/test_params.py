"""Pins of the oracle's parameter override (CPU).

The differentiable parameters theta (REBO pair parameters and LJ epsilon,
sigma per type pair) are set in the oracle the way a harness subclass of the
reference would set its protected members.  Each group is pinned by a closed
form of the energy at perturbed parameters on a geometry where every other
factor is known: perfect diamond (b_ij in closed form), ethane with the
attractive terms switched off, H2, and isolated atom pairs for the LJ term.
"""
import math

import numpy as np
import pytest

from mdtk_amd import systems as S
from tests.helpers import EV, ANG

REBO, AIREBO = 4, 8
# REBO.cxx:856-941 (eV, Angstrom), per pair CC, CH, HH: Q, A, alpha, B1-3, beta1-3
REF = {
    0: (0.313460, 10953.544, 4.7465391, 12388.792, 17.567065, 30.714932, 4.7204523, 1.4332132, 1.3826913),
    1: (0.340776, 149.94099, 4.1025498, 32.355187, 0.0, 0.0, 1.4344581, 0.0, 0.0),
    2: (0.370471, 32.817356, 3.5362986, 29.632593, 0.0, 0.0, 1.7158922, 0.0, 0.0),
}
UNITS = (ANG, EV, 1 / ANG, EV, EV, EV, 1 / ANG, 1 / ANG, 1 / ANG)


def _oracle(sys, pots):
    from oracle.oracle import Oracle
    return Oracle.from_system(sys, pots=pots)


def _pair_energy(r, p):
    """VR and VA of REBO.cxx:132-198 for cutoff 1, in eV (r in Angstrom)."""
    Q, A, al, B1, B2, B3, b1, b2, b3 = p
    VR = (1 + Q / r) * A * math.exp(-al * r)
    VA = -(B1 * math.exp(-b1 * r) + B2 * math.exp(-b2 * r) + B3 * math.exp(-b3 * r))
    return VR, VA


def _perturbed(seed):
    from mdtk_amd.systems import uniforms
    u = uniforms(99, 9000 + seed)
    out = {}
    for p in range(3):
        out[p] = tuple(v * (1 + 0.02 * (u[9 * p + k] - 0.5)) for k, v in enumerate(REF[p]))
    return out


def _theta(pp, eps=(0.00284, math.sqrt(0.00284 * 0.00150), 0.00150), sig=(3.40, 3.025, 2.65)):
    th = np.zeros(33)
    for p in range(3):
        th[9 * p:9 * p + 9] = np.array(pp[p]) * np.array(UNITS)
        th[27 + 2 * p] = eps[p] * EV
        th[28 + 2 * p] = sig[p] * ANG
    return th


def test_default_params_are_the_reference_constants(oracle_lib):
    th = _oracle(S.ethane(), REBO | AIREBO).params()
    assert np.allclose(th, _theta(REF), rtol=1e-15, atol=0)


def test_diamond_cc_params_closed_form(oracle_lib):
    """Perfect diamond: b_ij = (1 + 3 G_C2(-1/3))^(-1/2) does not depend on the
    pair parameters, so E/atom = 2 (VR + b VA) at any theta_CC."""
    pp = _perturbed(1)
    r = 3.567 * math.sqrt(3) / 4
    VR, VA = _pair_energy(r, pp[0])
    b = (1 + 3 * 0.097321) ** -0.5
    d = S.diamond(2, 2, 2)
    o = _oracle(d, REBO)
    o.set_params(_theta(pp))
    e, _ = o.energy_grad()
    assert abs(e[0] / EV / d.n - 2 * (VR + b * VA)) < 1e-9


def test_ethane_repulsive_params_closed_form(oracle_lib):
    """Ethane with B1-3 = 0 for every pair: VA = 0, and only the C-C and six
    C-H bonds are inside R2 (geminal H-H 1.78 A > 1.7 A), so
    E = VR_CC(1.54) + 6 VR_CH(1.09) at any Q, A, alpha."""
    pp = _perturbed(2)
    for p in range(3):
        pp[p] = pp[p][:3] + (0.0, 0.0, 0.0) + pp[p][6:]
    o = _oracle(S.ethane(), REBO)
    o.set_params(_theta(pp))
    e, _ = o.energy_grad()
    want = _pair_energy(1.54, pp[0])[0] + 6 * _pair_energy(1.09, pp[1])[0]
    assert abs(e[0] / EV - want) < 1e-9


def test_h2_params_closed_form(oracle_lib):
    """H2 at 0.74 A: no neighbours, so p_ij = p_ji = 1 and
    b = 1 + pi_rc_HH(0, 0, 1); E = VR + b VA at any theta_HH."""
    from oracle.oracle import spline_eval
    pp = _perturbed(3)
    sys = S.System(np.array([S.H, S.H], np.int32), np.array([[20.0, 20, 20], [20.74, 20, 20]]),
                   np.array([S.NO_PBC] * 3))
    o = _oracle(sys, REBO)
    o.set_params(_theta(pp))
    e, _ = o.energy_grad()
    VR, VA = _pair_energy(0.74, pp[2])
    b = 1 + spline_eval(7, 0.0, 0.0, 1.0)
    assert abs(e[0] / EV - (VR + b * VA)) < 1e-9


def _smooth_off(t):
    return 1 - t * t * (3 - 2 * t)


@pytest.mark.parametrize("pair,r", [((S.C, S.C), 4.3), ((S.C, S.H), 3.9), ((S.H, S.H), 3.4),
                                    ((S.C, S.C), 3.75), ((S.H, S.H), 3.2)])
def test_lj_params_closed_form(oracle_lib, pair, r):
    """Two isolated atoms: C_ij = 1 (no paths), the REBO term is 0 (r > R2)
    and f_LJ = 1 (r < R0).  For r above 2^(1/6) sigma, S_r = 0 and
    E = 4 eps ((sigma/r)^12 - (sigma/r)^6); inside [sigma, 2^(1/6) sigma],
    E = (1 + S_r (S_b(b*) - 1)) V with b* = 1 + pi_rc(0,0,1) of the pair."""
    from oracle.oracle import spline_eval
    eps = (0.0031, 0.0024, 0.0016)
    sig = (3.52, 3.10, 2.83)
    pp = _perturbed(4)
    sys = S.System(np.array(pair, np.int32), np.array([[20.0, 20, 20], [20.0 + r, 20, 20]]),
                   np.array([S.NO_PBC] * 3))
    o = _oracle(sys, REBO | AIREBO)
    o.set_params(_theta(pp, eps, sig))
    e, _ = o.energy_grad()
    p = (pair[0] == S.H) + (pair[1] == S.H)
    s6 = (sig[p] / r) ** 6
    V = 4 * eps[p] * (s6 * s6 - s6)
    rmax = sig[p] * 2 ** (1 / 6)
    if r >= rmax:
        want = V
    else:
        t = (r - sig[p]) / (rmax - sig[p])
        Sr = _smooth_off(t)
        bstar = 1 + spline_eval(5 + p, 0.0, 0.0, 1.0)
        bmin, bmax = [(0.77, 0.81), (0.75, 0.90), (0.32, 0.42)][p]
        tb = min(max((bstar - bmin) / (bmax - bmin), 0.0), 1.0)
        Sb = _smooth_off(tb)
        want = (1 + Sr * (Sb - 1)) * V
    assert abs(e[0]) < 1e-20
    assert abs(e[1] / EV - want) < 1e-12
