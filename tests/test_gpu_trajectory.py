"""Trajectory output captured on the GPU against the oracle stepped one
iteration at a time, with the cadence evaluated here on the oracle's clock:
SnapshotList records of the projectile at mdtrajsim's intervals
(mdtrajsim/main.cpp:58-83) and the XVA frame of iteration 0
(SimLoop.cxx:271-277), for a batch of two trajectories on clocks of their own."""
import numpy as np
import pytest

from mdtk_amd import systems as S
from mdtk_amd.engine import FRAME_SNAPSHOT, FRAME_XVA

pytestmark = pytest.mark.gpu

POTS = ("FBZL", "REBO", "AIREBO", "ETors")
STEPS = 260


def _due(t, dt_prev, interval):
    return t == 0.0 or int(t / interval) != int((t - dt_prev) / interval)


def _oracle_snapshots(s):
    from oracle.oracle import Oracle, POT_FBZL, POT_REBO, POT_AIREBO, POT_ETORS
    e, x, v, pbc, fx = s.to_cgs()
    o = Oracle(e, x, v, pbc, fx, pots=POT_FBZL | POT_REBO | POT_AIREBO | POT_ETORS)
    t, dtp, shots, xva = 0.0, 1e-20, [], []
    for it in range(STEPS):
        isnap = 1e-14 if t < 2e-12 else 5e-14 if t < 4e-12 else 1e-13
        xs, vs = o.state()[:2]
        if _due(t, dtp, isnap):
            shots.append((it, t, xs[-1].copy(), vs[-1].copy()))
        if it == 0 or _due(t, dtp, 1e-12 if t < 2e-12 else 5e-12):
            xva.append((it, t, xs.copy(), vs.copy()))
        dtp = o.run(1)["dt"][0]
        t += dtp
    return shots, xva


def test_snapshots_and_xva_match_oracle():
    from mdtk_amd import Engine
    mem = [S.bombard_target(300, 51, energy_ev=80.0, height=2.0), S.bombard_target(300, 52, energy_ev=40.0, height=3.0)]
    eng = Engine.from_systems(mem, potentials=POTS)
    sel = np.concatenate([np.eye(1, m.n, m.n - 1, dtype=np.uint8)[0] for m in mem])  # the projectiles
    eng.set_trajectory_output(sel, xva=True)
    eng.run(STEPS)
    fr, rec = eng.take_trajectory()
    for s, m in enumerate(mem):
        shots, xva = _oracle_snapshots(m)
        g = fr[(fr["system"] == s) & (fr["kind"] == FRAME_SNAPSHOT)]
        print(f"system {s}: {len(shots)} oracle snapshots, {len(g)} GPU", flush=True)
        assert len(g) == len(shots) and len(shots) > 3
        for f, (it, t, xo, vo) in zip(g, shots):
            assert f["iteration"] == it and abs(f["sim_time"] - t) <= 1e-12 * max(t, 1e-30)
            r = rec[f["first"]:f["first"] + f["count"]]
            assert len(r) == 1 and r["atom"][0] == m.n - 1
            assert np.abs(r["x"][0] - xo).max() <= 1e-9 * 1e-8
            assert np.abs(r["v"][0] - vo).max() <= 1e-9 * np.abs(vo).max()
        gx = fr[(fr["system"] == s) & (fr["kind"] == FRAME_XVA)]
        assert len(gx) == len(xva) == 1 and gx[0]["iteration"] == 0
        r = rec[gx[0]["first"]:gx[0]["first"] + gx[0]["count"]]
        assert len(r) == m.n and np.array_equal(r["atom"], np.arange(m.n))
        assert np.abs(r["x"] - xva[0][2]).max() <= 1e-9 * 1e-8
    # nothing is kept once taken; output off leaves the step as before
    assert len(eng.take_trajectory()[0]) == 0
    eng.set_trajectory_output(None, False)
    eng.run(5)
    assert len(eng.take_trajectory()[0]) == 0
