"""FBM (Born-Mayer pairs, FBM.cxx) and TightBinding (Cu second moment,
TightBinding.cxx) of the oracle against what the physics fixes (CPU).

FBM: a Cu-Cu or Cu-Ar pair at r is A exp(-b r) times the cosine switch
3.1-3.5 A (FPairwise's default range), with the Born-Mayer constants of
FBM.cxx:43-52 (Cu-Cu 22.565 keV, 5.088/A; Cu-Ar 59.874 keV, 7.2/A); Ar-Ar and
H/C pairs are not handled.

TightBinding: E = sum_i -c sqrt(rho_i) + sum_{i<j} V.  The dimer and the
equilateral triangle separate rho's pair function g and V at any r, without
retyping the parameters:
* g is one exponential (log-linear in r) inside 5.0 A and falls to half of
  that exponential's value at 5.25 A (the cosine switch's midpoint), 0 from
  5.5 A;
* V below 1.0 A is FBM's Cu-Cu Born-Mayer pair (the two reference classes
  share those constants), V is C1 across the 1.0-1.2 A bridge and one
  exponential beyond it, switched like g;
* the fcc crystal sits at zero pressure at Cu's 3.615 A and binds by Cu's
  cohesive energy, 3.49 eV;
* forces are -dE/dx (central differences).
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from oracle.oracle import Oracle, POT_FBM, POT_TB

EV = S.EV
ANG = S.ANG


def energy(elem, pos_a, pots, pbc=None):
    pbc = np.array([S.NO_PBC] * 3) if pbc is None else pbc
    s = S.System(np.asarray(elem, np.int32), np.asarray(pos_a, float) + 20.0 * (pbc == S.NO_PBC), pbc, name="x")
    e, g = Oracle.from_system(s, pots=pots).energy_grad()
    return float(e.sum()), g


def switch(r):
    return 1.0 if r < 3.1 else 0.0 if r > 3.5 else 0.5 * (1 + np.cos(np.pi * (r - 3.1) / 0.4))


@pytest.mark.parametrize("other,A_kev,b", [(64, 22.565, 5.088), (40, 59.874, 7.2)])
def test_fbm_pairs_are_born_mayer(other, A_kev, b):
    for r in (0.5, 1.5, 2.5, 3.0, 3.3, 3.45):
        e = energy([64, other], [[0, 0, 0], [r, 0, 0]], POT_FBM)[0]
        want = A_kev * 1000 * EV * np.exp(-b * r) * switch(r)
        assert abs(e - want) <= 1e-12 * abs(want)
    assert energy([64, other], [[0, 0, 0], [3.51, 0, 0]], POT_FBM)[0] == 0.0


def test_fbm_other_pairs_and_forces():
    for e1, e2 in ((40, 40), (64, 12), (40, 1), (108, 64)):
        e, g = energy([e1, e2], [[0, 0, 0], [2.0, 0, 0]], POT_FBM)
        assert e == 0.0 and not g.any()
    rng = np.random.default_rng(4)
    p = S.fcc(64, 2, 2, 2, a=3.0, periodic=(False, False, False)).pos + rng.normal(scale=0.15, size=(32, 3))
    el = np.where(rng.random(32) < 0.3, 40, 64)
    _, g = energy(el, p, POT_FBM)
    _fd(el, p, POT_FBM, g)


def _fd(el, p, pots, g, atoms=(0, 7, 19, 31), h=1e-5):
    for i in atoms:
        for d in range(3):
            q = p.copy()
            q[i, d] += h
            ep = energy(el, q, pots)[0]
            q[i, d] -= 2 * h
            em = energy(el, q, pots)[0]
            assert abs(g[i, d] - (ep - em) / (2 * h * ANG)) <= 1e-6 * np.abs(g).max()


def tb_terms(r):
    """(c^2 g, V) of Cu at distance r (A), from the dimer and the triangle."""
    e2 = energy([64, 64], [[0, 0, 0], [r, 0, 0]], POT_TB)[0]
    e3 = energy([64] * 3, [[0, 0, 0], [r, 0, 0], [r / 2, r * np.sqrt(3) / 2, 0]], POT_TB)[0]
    # dimer: -2 c sqrt(g) + V; triangle: -3 c sqrt(2 g) + 3 V; q = c sqrt(g)
    q = (e3 - 3 * e2) / (6 - 3 * np.sqrt(2))
    return q * q, e2 + 2 * q


def _loglinear(f, rs):
    y = np.log([f(r) for r in rs])
    slope = (y[-1] - y[0]) / (rs[-1] - rs[0])
    return slope, y[0] - slope * rs[0], y


def test_tb_density_is_one_switched_exponential():
    rs = np.array([1.5, 2.5, 3.5, 4.5, 4.9])
    slope, icpt, y = _loglinear(lambda r: tb_terms(r)[0], rs)
    assert np.abs(y - (icpt + slope * rs)).max() < 1e-9
    assert abs(tb_terms(5.25)[0] / np.exp(icpt + slope * 5.25) - 0.5) < 1e-9
    assert tb_terms(5.5 + 1e-6)[0] == pytest.approx(0.0, abs=1e-60)


def test_tb_pair_term_short_range_bridge_and_tail():
    for r in (0.5, 0.8, 0.95):
        v = tb_terms(r)[1]
        fbm = energy([64, 64], [[0, 0, 0], [r, 0, 0]], POT_FBM)[0]
        assert abs(v / fbm - 1) < 1e-9
    h = 1e-6
    for x in (1.0, 1.2):
        vm = [tb_terms(x - k * h)[1] for k in (2, 1)]
        vp = [tb_terms(x + k * h)[1] for k in (1, 2)]
        s_in, s_out = (vm[1] - vm[0]) / h, (vp[1] - vp[0]) / h
        assert abs(s_in / s_out - 1) < 1e-3  # slope continuous
        assert abs(vp[0] - vm[1] - h * (s_in + s_out)) <= 1e-8 * abs(vm[1])  # no jump
    rs = np.array([1.3, 2.0, 3.0, 4.0, 4.9])
    slope, icpt, y = _loglinear(lambda r: tb_terms(r)[1], rs)
    assert np.abs(y - (icpt + slope * rs)).max() < 1e-7
    assert abs(tb_terms(5.25)[1] / np.exp(icpt + slope * 5.25) - 0.5) < 1e-7


def test_tb_crystal_binds_like_copper():
    def e_atom(a):
        c = S.fcc(64, 6, 6, 6, a=a)
        return energy(c.elem, c.pos, POT_TB, pbc=c.pbc)[0] / c.n
    a, h = 3.615, 1e-4
    e0, ep, em = e_atom(a), e_atom(a * (1 + h)), e_atom(a * (1 - h))
    a0 = a - (ep - em) / (2 * h * a) / ((ep - 2 * e0 + em) / (h * a) ** 2)
    assert abs(a0 - 3.615) < 0.002
    assert abs(e0 / EV + 3.49) < 0.01


def test_tb_forces_and_other_elements():
    rng = np.random.default_rng(6)
    p = S.fcc(64, 2, 2, 2, a=3.615, periodic=(False, False, False)).pos + rng.normal(scale=0.15, size=(32, 3))
    el = np.full(32, 64)
    e0, g = energy(el, p, POT_TB)
    _fd(el, p, POT_TB, g)
    # Ag, Au, C and Ar atoms among them change nothing
    q = np.vstack([p, p[3] + [0, 0, 2.0], p[9] + [2.0, 0, 0], p[12] + [0, 2.0, 0], p[20] + [0, -2.0, 0]])
    e1, g1 = energy(np.concatenate([el, [108, 197, 12, 40]]), q, POT_TB)
    assert e1 == e0 and np.array_equal(g1[:32], g) and not g1[32:].any()
