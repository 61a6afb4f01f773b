"""The oracle against the reference's own code where that code builds here.

oracle/_ref/libmdtkref.so is compiled in place from the reference tree's
GSL-free translation units (oracle/ref.mk): Atom/AtomsArray, NeighbourList,
FGeneral's CosTheta/CosDihedral, FProxy, FBZL and FBM.  These tests run the
same inputs through it and through the oracle's restatement (CPU only).  They
are skipped where the library was not built (no reference tree).
"""
import numpy as np
import pytest

from mdtk_amd import systems as S
from oracle import oracle as O
from oracle import ref

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (no reference tree)")

A = 1e-8
NO = S.NO_PBC * A
H, C, AR, CU, XE, AU, AG = 1, 12, 40, 64, 131, 197, 108


def _rng(seed):
    return np.random.default_rng(seed)


def test_depos_same_and_mixed_pbc():
    """depos (Atom.hpp:115-140): one +-L shift for a shared period, the
    PBC_count unwrap when only one atom of the pair has the axis periodic,
    nothing for open axes -- bit for bit."""
    rng = _rng(1)
    L = np.array([20.0, 17.0, 30.0]) * A
    for _ in range(400):
        xa, xb = rng.uniform(-5, 35, 3) * A, rng.uniform(-5, 35, 3) * A
        ca, cb = rng.integers(-3, 4, 3), rng.integers(-3, 4, 3)
        La, Lb = L.copy(), L.copy()
        for d in range(3):  # per axis: both periodic, one open, or both open
            k = rng.integers(0, 4)
            if k == 1:
                La[d] = NO
            elif k == 2:
                Lb[d] = NO
            elif k == 3:
                La[d] = Lb[d] = NO
        ca = np.where(La == NO, 0, ca)  # Atom::applyPBC requires PBC_count == 0 on an open axis
        cb = np.where(Lb == NO, 0, cb)
        r_ref = ref.depos(xa, La, ca, xb, Lb, cb)
        r_orc = O.depos(xa, La, ca, xb, Lb, cb)
        assert np.array_equal(r_ref, r_orc), (xa, La, ca, xb, Lb, cb)


def test_apply_pbc():
    """Atom::applyPBC (Atom.cxx:93-125): wrap into [0, L) and count the
    crossings, open axes untouched."""
    rng = _rng(2)
    L = np.array([12.5, 9.0, NO / A]) * A
    for _ in range(200):
        x = rng.uniform(-40, 50, 3) * A
        c = rng.integers(-2, 3, 3)
        c[2] = 0
        xr, cr = ref.apply_pbc(x, L, c)
        xo, co = O.apply_pbc(x, L, c)
        assert np.array_equal(xr, xo) and np.array_equal(cr, co)


def test_cos_theta_and_adjoint():
    """CosTheta and its seeded adjoint (FGeneral.hpp:138-185); the oracle's
    vector-form adjoint equals the reference's to rounding."""
    rng = _rng(3)
    for _ in range(200):
        x = rng.normal(0, 1.2, (3, 3)) * A
        V = rng.normal() * 1e-12
        cr, gr = ref.cos_theta(x, V)
        co, go = O.cos_theta(x, V)
        assert abs(cr - co) <= 1e-15
        assert np.abs(gr - go).max() <= 1e-12 * np.abs(gr).max()


def test_cos_dihedral_and_adjoint():
    """CosDihedral and its hand-expanded adjoint (FGeneral.hpp:201-305)
    against the oracle's vector-form adjoint, on random quads and on planar
    (cis, trans) and near-collinear ones."""
    rng = _rng(4)
    quads = [rng.normal(0, 1.3, (4, 3)) * A for _ in range(200)]
    base = np.array([[0, 0, 0], [1.5, 0, 0], [-0.5, 1.4, 0], [2.0, 1.4, 0]]) * A
    quads.append(base)                                   # cis, planar
    trans = base.copy()
    trans[3, 1] *= -1
    quads.append(trans)                                  # trans, planar
    tw = base.copy()
    tw[3] = np.array([2.0, 1.4 * np.cos(1.0), 1.4 * np.sin(1.0)]) * A
    quads.append(tw)
    for x in quads:
        V = 1e-12
        cr, gr = ref.cos_dihedral(x, V)
        co, go = O.cos_dihedral(x, V)
        assert abs(cr - co) <= 2e-15
        assert np.abs(gr - go).max() <= 1e-11 * max(np.abs(gr).max(), 1e-300)


def _ion_cluster(seed, n=40, ion=AR, targets=(H, C)):
    rng = _rng(seed)
    pos = rng.uniform(0, 9, (n, 3))
    # keep atoms at least 0.9 A apart (ZBL is steep but finite)
    keep = [0]
    for i in range(1, n):
        if np.min(np.linalg.norm(pos[keep] - pos[i], axis=1)) > 0.9:
            keep.append(i)
    pos = pos[keep]
    el = rng.choice(list(targets), len(pos))
    el[:3] = ion
    return el.astype(np.int32), pos * A


@pytest.mark.parametrize("seed", [5, 6, 7])
def test_fbzl_energy_and_gradient(seed):
    """FBZL (pairwise/FBZL.cxx:70-127) with common.h's 5.0-5.5 A switch:
    energies and gradients of the oracle's restatement equal the reference
    code's on Ar/Xe ions among H, C, Cu, Ag, Au."""
    el, x = _ion_cluster(seed, targets=(H, C, CU, AG, AU))
    el[3] = XE
    pbc = np.array([NO, NO, NO])
    er, gr = ref.pair_chain(el, x, pbc, 1)
    orc = O.Oracle(el, x, None, pbc, pots=O.POT_FBZL)
    eo, go = orc.energy_grad()
    assert er[0] != 0.0
    assert abs(eo[0] - er[0]) <= 1e-14 * abs(er[0])
    assert np.abs(go - gr).max() <= 1e-13 * np.abs(gr).max()


@pytest.mark.parametrize("seed", [8, 9])
def test_fbm_energy_and_gradient(seed):
    """FBM (pairwise/FBM.cxx:34-94, default 3.1-3.5 A cutoff) on Cu/Ar
    clusters, periodic in x."""
    el, x = _ion_cluster(seed, n=50, ion=AR, targets=(CU,))
    pbc = np.array([9.5 * A, NO, NO])
    er, gr = ref.pair_chain(el, x, pbc, 128)
    orc = O.Oracle(el, x, None, pbc, pots=O.POT_FBM)
    eo, go = orc.energy_grad()
    assert er[0] != 0.0
    assert abs(eo[0] - er[0]) <= 1e-14 * abs(er[0])
    assert np.abs(go - gr).max() <= 1e-13 * np.abs(gr).max()


def _member_lists(sys_, pots, kind):
    e, x, v, p, f = sys_.to_cgs()
    orc = O.Oracle(e, x, v, p, f, pots=pots)
    orc.energy_grad()
    return e, x, p, orc.neighbour_lists(orc.kinds.index(kind))


@pytest.mark.parametrize("case", ["pe_periodic", "slab_ion", "diamond"])
def test_neighbour_lists_match_reference(case):
    """NeighbourList::Update (NeighbourList.cxx:31-106): the oracle's REBO
    (2.0 A, H/C), AIREBO (5.5 A, H/C) and FBZL (5.5 A, its element set) lists
    equal the reference's, entry for entry and in order."""
    if case == "pe_periodic":
        s = S.jitter(S.polyethylene(2, 3, 5), 0.1, 3)
    elif case == "slab_ion":
        s = S.bombard_target(600, 1)
    else:
        s = S.jitter(S.diamond(4, 4, 4), 0.08, 4)
    for kind, rc, handled in ((O.POT_REBO, 2.0, (H, C)), (O.POT_AIREBO, 5.5, (H, C)),
                              (O.POT_FBZL, 5.5, (H, C, CU, AU, AG, AR, XE))):
        pots = O.POT_REBO | O.POT_AIREBO | O.POT_FBZL
        e, x, p, lo = _member_lists(s, pots, kind)
        lr = ref.neighbour_list(e, x, p, rc * A, handled)
        assert len(lo) == len(lr)
        for i in range(len(lo)):
            assert np.array_equal(lo[i], lr[i]), (case, kind, i)


def test_skin_rule_matches_reference():
    """MovedTooMuch (NeighbourList.cxx:108-133): rebuild when the two largest
    accumulated displacements sum past 0.44 Rc, at and around the threshold."""
    rng = _rng(10)
    rc = 2.0 * A
    for _ in range(300):
        n = int(rng.integers(2, 30))
        d = rng.normal(0, 0.2, (n, 3)) * A
        # scale so that the top-two sum lands near the threshold
        m = np.sort(np.linalg.norm(d, axis=1))[-2:].sum()
        d *= (0.44 * rc / m) * rng.choice([0.999999, 1.0, 1.000001, 0.5, 2.0])
        assert ref.moved_too_much(d, rc) == O.moved_too_much(d, rc)


def test_maxwell_momentum_removal_matches_reference():
    """The synthetic-input generator's zero-momentum step equals
    AtomsArray::removeMomentum (AtomsContainer.cxx:205-219,268-278): the
    mass-weighted mean velocity of the non-fixed atoms is subtracted, fixed
    atoms stand still."""
    s = S.bombard_target(700, 3)
    m = np.array([S.MASS_AMU[int(e)] for e in s.elem]) * S.AMU
    z = S.normals(3 * s.n, 3).reshape(s.n, 3)
    raw = z * np.sqrt(S.KB * 300.0 / m)[:, None]  # cm/s, before the removal
    vr = ref.remove_momentum(s.elem, raw, s.fixed)
    vo = S.maxwell_velocities(s, 300.0, stream=3) * (S.ANG / S.FS)
    assert np.abs(vr - vo).max() <= 1e-12 * np.abs(vr).max()
