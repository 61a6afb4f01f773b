"""An engine started from a state in the SimLoopSaver layout runs as one
started from the same arrays, bath included, and its state written back
reads as the engine's own (GPU)."""
import numpy as np
import pytest

from mdtk_amd import state as ST
from mdtk_amd import systems as S

pytestmark = pytest.mark.gpu

POTS = ("REBO", "AIREBO", "ETors")


def test_engine_from_state_runs_like_direct(tmp_path):
    from mdtk_amd import Engine
    s = S.pe_slab(2)
    s.vel = S.maxwell_velocities(s, 500.0, stream=4)
    e, x, v, pbc, fx = s.to_cgs()
    n = s.n
    top = s.pos[:, 2].max()
    bath = dict(geometry=2, To=300.0, gamma=2.0e13, z_min=(top - 3.0) * S.ANG, d_boundary=2.0 * S.ANG,
                z_min_free=-5.0 * S.ANG, center=(0.0, 0.0, 0.0), radius=1e-2)
    st = ST.State(e, x.reshape(n, 3), v.reshape(n, 3), np.zeros((n, 3)), pbc, np.zeros((n, 3), np.int32),
                  np.tile(pbc != ST.NO_PBC, (n, 1)), np.arange(n, dtype=np.uint32), {"fixed": fx.astype(bool)},
                  bath)
    p = str(tmp_path / "base")
    ST.write_state(p, st)
    a = ST.engine_from_state(ST.read_state(p), potentials=POTS)
    b = Engine.from_system(s, potentials=POTS)
    b.set_thermal_bath(**{k: bath[k] for k in ("To", "gamma", "z_min", "d_boundary", "z_min_free")}, geometry="box")
    ra = a.run(30, fixed_dt=0.2e-15)
    rb = b.run(30, fixed_dt=0.2e-15)
    assert np.abs(ra["epot"] - rb["epot"]).max() <= 1e-12 * np.abs(rb["epot"]).max()
    assert np.abs(ra["ekin"] - rb["ekin"]).max() <= 1e-12 * np.abs(rb["ekin"]).max()
    assert abs(a.bath_energy()[0] - b.bath_energy()[0]) <= 1e-10 * abs(b.bath_energy()[0])
    out = ST.update_state(st, a, sim_time_advance=30 * 0.2e-15, dt=0.2e-15)
    ST.write_state(p + "_after", out, compress="gz")
    back = ST.read_state(p + "_after")
    xa, va, aa, ca = a.state()
    assert np.array_equal(back.pos, xa) and np.array_equal(back.vel, va) and np.array_equal(back.acc, aa)
    assert np.array_equal(back.pbc_count, ca)
    assert abs(back.time["sim_time"] - 30 * 0.2e-15) < 1e-25
