"""Synthetic input generators (CPU): geometry and seeding of the workloads."""
import numpy as np

from mdtk_amd import systems as S


def _mic_dist(p, q, pbc):
    d = q - p
    for k in range(3):
        if pbc[k] != S.NO_PBC:
            d[..., k] -= pbc[k] * np.round(d[..., k] / pbc[k])
    return np.sqrt((d * d).sum(-1))


def test_pe_slab_bonding():
    s = S.pe_slab(3)
    assert s.n == 12 * 8 * 3 * 3
    assert s.pbc[2] == S.NO_PBC
    Cp, Hp = s.pos[s.elem == S.C], s.pos[s.elem == S.H]
    for h in Hp:
        assert abs(_mic_dist(h, Cp, s.pbc).min() - 1.09) < 1e-9
    for c in Cp:
        d = np.sort(_mic_dist(c, Cp, s.pbc))
        assert abs(d[1] - 1.54) < 1e-9 and abs(d[2] - 1.54) < 1e-9 and d[3] > 2.4


def test_bombard_target():
    t = S.bombard_target(5000, 3)
    assert t.elem[-1] == S.AR and t.fixed[-1] == 0
    n_layers = t.meta["layers"]
    assert t.n == 288 * n_layers + 1
    # bottom chain row fixed and at rest; ion above the surface moving down at 100 eV
    assert t.fixed.sum() == 6 * 8 * 3
    assert np.all(t.vel[t.fixed == 1] == 0)
    assert t.pos[-1, 2] > t.pos[:-1, 2].max() + 7.9
    m = S.MASS_AMU[S.AR] * S.AMU
    v = t.vel[-1] * (S.ANG / S.FS)
    assert abs(0.5 * m * (v ** 2).sum() / S.EV - 100.0) < 1e-9
    assert v[0] == 0 and v[1] == 0 and v[2] < 0
    # zero net momentum of the free slab atoms
    slab = (t.fixed == 0) & (t.elem != S.AR)
    ms = np.array([S.MASS_AMU[int(e)] for e in t.elem[slab]])
    assert np.abs((t.vel[slab] * ms[:, None]).sum(0)).max() < 1e-9 * np.abs(t.vel[slab]).max() * ms.sum()


def test_bombard_batch_seeded_sizes():
    a = S.bombard_batch(16)
    b = S.bombard_batch(16)
    assert [x.n for x in a] == [x.n for x in b]
    assert all(np.array_equal(x.pos, y.pos) for x, y in zip(a, b))
    assert len({x.n for x in a}) > 4
    assert all(2000 - 288 <= x.n - 1 <= 12000 + 288 for x in a)
