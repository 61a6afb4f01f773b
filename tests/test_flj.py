"""FLJ (hydrocarbon-metal pairs, FLJ.cxx) of the oracle against textbook forms (CPU).

A pair (C or H) x (Cu or Au) at distance r interacts by
* the Ziegler-Biersack-Littmark universal potential below 1.2 A:
  Z1 Z2 e^2 / r * phi(r / a), phi(x) = 0.18175 e^-3.1998x + 0.50986 e^-0.94229x
  + 0.28022 e^-0.4029x + 0.02817 e^-0.20162x, a = 0.88534 a0 / (Z1^0.23 + Z2^0.23),
  with the reference's a0 = 0.53 A and e = 4.8e-10 esu;
* Lennard-Jones beyond 1.9 A, 4 eps ((s/r)^12 - (s/r)^6) with Lorentz-Berthelot
  mixing of the like-pair values (Cu 3.05 A / 0.165565 eV, C 3.4 A / 0.002413
  eV, Au 2.569 A / 0.408 eV, H 2.81 A / 8.6 K), times the cosine switch from
  5.0 to 5.5 A;
* a cubic between 1.2 and 1.9 A matching value and slope at both ends.
Other pairs (C-H, metal-metal) get nothing.  Forces are -dE/dx.
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from oracle.oracle import Oracle, POT_FLJ

EV = S.EV
ANG = S.ANG
E_ESU = 4.8e-10
KB = 8.617343e-5 * EV
LIKE = {12: (3.4, 0.002413 * EV), 1: (2.81, 8.6 * KB), 64: (3.05, 0.165565 * EV), 197: (2.569, 0.408 * EV)}
ZNUM = {1: 1, 12: 6, 64: 29, 197: 79}


def zbl(z1, z2, r_cm):
    a = 0.88534 * 0.53e-8 / (z1 ** 0.23 + z2 ** 0.23)
    x = r_cm / a
    phi = 0.18175 * np.exp(-3.1998 * x) + 0.50986 * np.exp(-0.94229 * x) + 0.28022 * np.exp(-0.4029 * x) \
        + 0.02817 * np.exp(-0.20162 * x)
    return (z1 * E_ESU) * (z2 * E_ESU) / r_cm * phi


def lj(e1, e2, r_cm):
    s = (LIKE[e1][0] + LIKE[e2][0]) / 2 * ANG
    eps = np.sqrt(LIKE[e1][1] * LIKE[e2][1])
    q = (s / r_cm) ** 6
    return 4 * eps * (q * q - q)


def switch(r_a):
    if r_a < 5.0:
        return 1.0
    if r_a > 5.5:
        return 0.0
    return 0.5 * (1 + np.cos(np.pi * (r_a - 5.0) / 0.5))


def pair(e1, e2, r_a):
    s = S.System(np.array([e1, e2], np.int32), np.array([[10.0, 10.0, 10.0], [10.0 + r_a, 10.0, 10.0]]),
                 np.array([S.NO_PBC] * 3), name="pair")
    o = Oracle.from_system(s, pots=POT_FLJ)
    e, g = o.energy_grad()
    return float(e.sum()), g


@pytest.mark.parametrize("e1,e2", [(12, 64), (1, 64), (12, 197), (1, 197), (64, 12)])
def test_regions_match_textbook_forms(e1, e2):
    z1, z2 = ZNUM[e1], ZNUM[e2]
    for r in (0.6, 0.9, 1.15):
        assert abs(pair(e1, e2, r)[0] / zbl(z1, z2, r * ANG) - 1) < 1e-12
    for r in (2.0, 3.0, 4.5, 5.2, 5.45):
        want = lj(e1, e2, r * ANG) * switch(r)
        assert abs(pair(e1, e2, r)[0] - want) <= 1e-12 * abs(want) + 1e-30
    assert pair(e1, e2, 5.6)[0] == 0.0


@pytest.mark.parametrize("e1,e2", [(12, 64), (1, 197)])
def test_bridge_is_c1_and_forces_are_gradients(e1, e2):
    z1, z2 = ZNUM[e1], ZNUM[e2]
    h = 1e-6
    # continuity of value and slope at both ends of the bridge
    for x, f in ((1.2, lambda r: zbl(z1, z2, r * ANG)), (1.9, lambda r: lj(e1, e2, r * ANG))):
        assert abs(pair(e1, e2, x + 1e-9)[0] / f(x) - 1) < 1e-6
        slope_in = (pair(e1, e2, x + 2 * h)[0] - pair(e1, e2, x + h)[0]) / h
        slope_out = (f(x - h) - f(x - 2 * h)) / h
        assert abs(slope_in / slope_out - 1) < 1e-4
    # force on atom 1 along x equals -dE/dr everywhere
    for r in (1.0, 1.3, 1.6, 1.85, 2.5, 5.25):
        e_p, e_m = pair(e1, e2, r + h)[0], pair(e1, e2, r - h)[0]
        dEdr = (e_p - e_m) / (2 * h * ANG)
        g = pair(e1, e2, r)[1]
        assert abs(g[1, 0] / dEdr - 1) < 1e-6  # atom 1 sits at larger x: dE/dx1 = dE/dr
        assert abs(g[0, 0] + g[1, 0]) <= 1e-12 * abs(g[1, 0])


def test_other_pairs_do_not_interact():
    for e1, e2 in ((12, 1), (64, 197), (12, 12), (64, 64)):
        e, g = pair(e1, e2, 2.5)
        assert e == 0.0 and not g.any()
