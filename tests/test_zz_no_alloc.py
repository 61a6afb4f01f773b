"""mdg_energy_grad_noalloc (no device allocation, static scratch) against
mdg_energy_grad on the same systems: two and more calls in a row on one GPU,
each smaller than the one before, so that a call meets the scratch the
previous one left behind."""
import numpy as np
import pytest

from mdtk_amd import Engine
from mdtk_amd import engine as E
from mdtk_amd import systems as S
from tests.helpers import force_rel_err, pe_small

CHAIN = ("FBZL", "REBO", "AIREBO", "ETors")


def _ref(sys_, pots, precision):
    e = Engine.from_system(sys_, potentials=pots, precision=precision)
    return e.energy_grad()


def _check(sys_, pots, precision="fp64"):
    print(f"noalloc: {sys_.name or 'system'} n={sys_.n} {precision}", flush=True)
    e0, g0 = _ref(sys_, pots, precision)
    e1, g1 = E.energy_grad_noalloc(sys_, potentials=pots, precision=precision)
    assert e1.shape == e0.shape
    assert np.abs(e1 - e0).max() <= 1e-12 * np.abs(e0).max()
    assert force_rel_err(g1, g0) <= 1e-12


@pytest.mark.gpu
def test_noalloc_matches_energy_grad_shrinking_calls():
    big = S.diamond(8, 8, 8)  # 4096 atoms: the limit
    big = S.jitter(big, 0.05, 2)
    _check(big, ("REBO", "AIREBO", "ETors"))
    _check(S.bombard_target(600, 4), CHAIN)
    _check(pe_small(), CHAIN)
    _check(pe_small(), CHAIN, "mixed")
    _check(S.hydrocarbon_cluster(60, 3), ("REBO", "AIREBO", "ETors"))


@pytest.mark.gpu
def test_noalloc_refuses_beyond_limit():
    s = S.diamond(8, 8, 9)  # 4608 atoms
    with pytest.raises(E.EngineError, match="4096"):
        E.energy_grad_noalloc(s, potentials=("REBO",))
