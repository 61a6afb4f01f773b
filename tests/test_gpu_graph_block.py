"""One graph launch per block of steps (mdg_set_graph_block): a batch stepped
32 steps per launch ends bit for bit where the same batch stepped one step
per launch ends (FP64 form: deterministic sums), with the same per-step
energies, dt and clocks, and a run whose length is not a multiple of the
block mixes both kinds of launch (GPU)."""
import numpy as np
import pytest

from mdtk_amd import systems as S

pytestmark = pytest.mark.gpu

POTS = ("FBZL", "REBO", "AIREBO", "ETors")


def _batch():
    return [S.bombard_target(n, k, height=3.0) for k, n in enumerate((600, 900, 300))]


@pytest.mark.parametrize("fixed_dt", [0.1e-15, 0.0])
def test_graph_block_equals_single_step_launches(fixed_dt):
    from mdtk_amd import Engine
    out = []
    for block in (1, 32, 7):
        e = Engine.from_systems(_batch(), potentials=POTS)
        e.set_graph_block(block)
        r = e.run(75, fixed_dt=fixed_dt)
        x, v, a, c = e.state()
        out.append((r, x, v, a, c, e.time()))
        print(f"block {block}: E_pot(last) = {r['epot'][-1]}", flush=True)
    r0, x0, v0, a0, c0, t0 = out[0]
    for r, x, v, a, c, t in out[1:]:
        for k in ("epot", "ekin", "dt"):
            assert np.array_equal(r[k], r0[k])
        assert np.array_equal(x, x0) and np.array_equal(v, v0) and np.array_equal(a, a0)
        assert np.array_equal(c, c0)
        assert np.array_equal(t, t0)
