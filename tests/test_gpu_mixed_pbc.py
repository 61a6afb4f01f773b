"""GPU parity for atoms that carry their own PBC vectors (Atom::PBC,
mdtk/Atom.hpp:115-140): a PE slab periodic in x and y with an Ar ion open on
every axis next to the x = 0 face, some slab atoms with nonzero PBC_count
(tests.helpers.mixed_pbc_case), against the oracle given the same per-atom
PBC vectors and counters.

FP64 bar as in test_gpu_parity: energies per member |dE| <= 1e-12 |E| (plus
a 1e-9 eV floor), forces max|dF| / max|F_ref| <= 1e-10, neighbour lists
identical; trajectories 1e-10 per step.  The mixed form's forces stay within
an RMS of 1e-5 eV/A of the oracle's."""
import numpy as np
import pytest

from mdtk_amd import systems as S
from mdtk_amd import state as ST
from tests.helpers import EV, EV_PER_A, force_rel_err, mixed_pbc_case

pytestmark = pytest.mark.gpu

POTS = ("FBZL", "REBO", "AIREBO", "ETors")
MASK = 1 | 4 | 8 | 16


def _pair(precision="fp64", lateral_v=0.0):
    from mdtk_amd import Engine
    from oracle.oracle import Oracle
    sys_, pbc, cnt = mixed_pbc_case()
    e, x, v, p, f = sys_.to_cgs()
    v = v.reshape(-1, 3).copy()
    v[-1, 0] = lateral_v
    o = Oracle(e, x, v, p, f, pots=MASK)
    o.set_atom_pbc(pbc, cnt)
    g = Engine(potentials=POTS, precision=precision)
    g.set_system(e, x, v, p, f)
    g.set_atom_pbc(pbc, x, cnt)
    print(f"mixed PBC case n={sys_.n} precision={precision}", flush=True)
    return sys_, o, g


def test_mixed_pbc_energy_force_lists():
    _, o, g = _pair()
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    print(f"  E_oracle(eV)={eo / EV} E_gpu(eV)={eg / EV}", flush=True)
    for a, b in zip(eo, eg):
        assert abs(a - b) <= 1e-12 * abs(a) + 1e-9 * EV, (a / EV, b / EV)
    fr = force_rel_err(gg, go)
    print(f"  force rel err {fr:.3e}", flush=True)
    assert fr <= 1e-10
    for term in range(4):
        lo, lg = o.neighbour_lists(term), g.neighbour_lists(term)
        for i, (a, b) in enumerate(zip(lo, lg)):
            assert np.array_equal(a, b), (term, i, a, b)


def test_mixed_pbc_trajectory():
    """The ion (100 eV down, 0.25 A/fs along -x) leaves the box through the
    x = 0 face and, being open, is not wrapped; slab atoms that cross a
    periodic face change their PBC_count, which its pairs with the ion then
    use.  40 steps of 0.1 fs."""
    sys_, o, g = _pair(lateral_v=-0.25 * S.ANG / S.FS)
    ro = o.run(40, fixed_dt=0.1e-15)
    rg = g.run(40, fixed_dt=0.1e-15)
    d = np.abs(rg["epot"] - ro["epot"]).max() / np.abs(ro["epot"]).max()
    print(f"  40 steps: max |dEpot|/|Epot| = {d:.2e}", flush=True)
    assert d <= 1e-10
    xo, vo, _, co = o.state()
    xg, vg, _, cg = g.state()
    assert xg[-1, 0] < 0.0  # the open ion left the box unwrapped
    assert np.array_equal(co, cg)
    assert np.abs(xg - xo).max() <= 1e-10 * S.ANG
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()


def test_mixed_pbc_mixed_precision_against_oracle():
    _, o, g = _pair(precision="mixed")
    eo, go = o.energy_grad()
    eg, gg = g.energy_grad()
    rms = np.sqrt((np.linalg.norm(gg - go, axis=1) ** 2).mean()) / EV_PER_A
    print(f"  mixed form: force RMS error {rms:.2e} eV/A", flush=True)
    assert rms <= 1e-5
    assert np.abs(eg - eo).max() / EV <= 1e-3


def test_mixed_pbc_batch_member_and_state(tmp_path):
    """A batch of a regular slab and the mixed frame: each member as alone;
    the mixed frame again from a SimLoopSaver state whose per-atom enabled
    flags leave the ion's axes open."""
    from mdtk_amd import Engine
    sys_, pbc, cnt = mixed_pbc_case()
    reg = S.bombard_target(288, 11, lateral=(8, 3))
    parts = [reg.to_cgs(), sys_.to_cgs()]
    b = Engine(potentials=POTS)
    b.set_systems(parts)
    e0, x0, v0, p0, f0 = parts[0]
    pbc_all = np.vstack([np.tile(p0, (reg.n, 1)), pbc])
    b.set_atom_pbc(pbc_all, np.concatenate([x0.reshape(-1), parts[1][1].reshape(-1)]),
                   np.vstack([np.zeros((reg.n, 3), np.int32), cnt]))
    eb, gb = b.energy_grad()
    _, _, g1 = _pair()
    e1, gg1 = g1.energy_grad()
    a0 = Engine.from_system(reg, potentials=POTS)
    ea0, ga0 = a0.energy_grad()
    assert np.abs(eb[1] - e1).max() <= 1e-12 * np.abs(e1).max()
    assert np.abs(eb[0] - ea0).max() <= 1e-12 * np.abs(ea0).max()
    assert force_rel_err(gb[reg.n:], gg1) <= 1e-12
    assert force_rel_err(gb[:reg.n], ga0) <= 1e-12
    # the same frame through a state file with per-atom PBC flags
    e, x, v, p, f = sys_.to_cgs()
    n = sys_.n
    en = pbc != ST.NO_PBC
    st = ST.State(e, x.reshape(n, 3), v.reshape(n, 3), np.zeros((n, 3)), p, cnt, en,
                  np.arange(n, dtype=np.uint32), {"fixed": f.astype(bool)}, None)
    path = str(tmp_path / "mixed")
    ST.write_state(path, st)
    back = ST.read_state(path)
    assert np.array_equal(back.pbc_enabled, en)
    es, gs = ST.engine_from_state(back, clock=False, potentials=POTS).energy_grad()
    assert np.abs(es - e1).max() <= 1e-12 * np.abs(e1).max()
    assert force_rel_err(gs, gg1) <= 1e-12
