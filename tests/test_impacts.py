"""The impact-sequence driver's file handling (CPU): ionpos.bin as raw
(x, y) float64 pairs in any of yaatk's compressions (main.cpp:291-301), the
projectile as the state's last atom (:306-308), the finished-directory
markers (:95-112) and the ionpos_used records (:318-327)."""
import os

import numpy as np
import pytest

from mdtk_amd import impacts as I
from mdtk_amd import state as ST


@pytest.mark.parametrize("compress", [None, "xz", "gz"])
def test_ionpos_round_trip(tmp_path, compress):
    xy = np.random.default_rng(1).random((7, 2)) * 3e-7
    p = str(tmp_path / "ionpos.bin")
    I.write_ionpos(p, xy, compress)
    assert np.array_equal(I.read_ionpos(p), xy)
    if compress is None:
        assert os.path.getsize(p) == 7 * 16


def test_ionpos_bad_length(tmp_path):
    p = tmp_path / "ionpos.bin"
    p.write_bytes(b"\0" * 24)
    with pytest.raises(ValueError):
        I.read_ionpos(str(p))


def test_projectile_is_the_last_atom_and_markers(tmp_path):
    n = 5
    st = ST.State(np.array([12, 1, 1, 12, 40], np.int32), np.arange(15, dtype=float).reshape(n, 3) * 1e-8,
                  np.zeros((n, 3)), np.zeros((n, 3)), np.array([1e-7, 1e-7, ST.NO_PBC]), np.zeros((n, 3), np.int32),
                  np.tile([True, True, False], (n, 1)), np.arange(n, dtype=np.uint32))
    s = I.impact_state(st, 1.5e-8, 2.5e-8)
    assert np.array_equal(s.pos[:-1], st.pos[:-1])
    assert s.pos[-1, 0] == 1.5e-8 and s.pos[-1, 1] == 2.5e-8 and s.pos[-1, 2] == st.pos[-1, 2]
    assert st.pos[-1, 0] == 12 * 1e-8  # the base state is untouched
    d = I.impact_dir(str(tmp_path), 42)
    assert os.path.basename(d) == "00000042"
    os.makedirs(d)
    assert not I.is_finished(d)
    for m in I.FINISHED:
        open(os.path.join(d, m), "w").close()
        assert I.is_finished(d)
        os.remove(os.path.join(d, m))
    I._write_used(d, 1.25e-7, 3e-8)
    assert open(os.path.join(d, "ionpos_used.txt")).read() == "1.25e-07 3e-08\n"
    assert np.array_equal(np.fromfile(os.path.join(d, "ionpos_used.bin"), "<f8"), [1.25e-7, 3e-8])


def _touch(d, names):
    for f in names:
        open(os.path.join(d, f), "w").close()


def test_iterations_listing_and_removal(tmp_path):
    """SimLoopSaver::listIterations / removeIterations(keepFirst, keepLast)
    (SimLoopSaver.cxx:718-795): ids md%010d; other files stay."""
    d = str(tmp_path)
    for it in (0, 1000, 2000, 2417):
        _touch(d, [I.state_id(it) + ext for ext in (".r.xz", ".v.xz", ".time.xz")])
    _touch(d, ["snapshots.conf", "mde0000001000.xva", "ionpos_used.bin"])
    assert I.state_id(2417) == "md0000002417"
    assert I.list_iterations(d) == [0, 1000, 2000, 2417]
    I.remove_iterations(d)  # runTraj: keep the first and the last
    assert I.list_iterations(d) == [0, 2417]
    I.remove_iterations(d, keep_first=False, keep_last=True)  # the impact loop
    assert I.list_iterations(d) == [2417]
    assert {"snapshots.conf", "mde0000001000.xva", "ionpos_used.bin"} <= set(os.listdir(d))


def _small_state(n=4):
    return ST.State(np.array([12, 1, 1, 40][:n], np.int32), np.arange(3 * n, dtype=float).reshape(n, 3) * 1e-8,
                    np.ones((n, 3)) * 1e4, np.ones((n, 3)) * -2e12, np.array([1e-7, 1e-7, ST.NO_PBC]),
                    np.zeros((n, 3), np.int32), np.tile([True, True, False], (n, 1)), np.arange(n, dtype=np.uint32),
                    {"fixed": np.array([1, 0, 0, 0][:n], bool), "projectile": np.array([0, 0, 0, 1][:n], bool)},
                    time=dict(sim_time=3e-14, sim_time_final=1e-12, dt=2e-16, dt_prev=1.9e-16))


def test_data_state_flag_during_writes(tmp_path, monkeypatch):
    """yaatk::DataState (yaatk.hpp:435-470): the flag exists while a state is
    written and is gone after; a flag left behind by an interrupted write is
    cleared by the next write; runTraj refuses a dirty directory
    (main.cpp:119-120)."""
    d = str(tmp_path)
    seen = []
    real = ST.write_state
    monkeypatch.setattr(ST, "write_state", lambda *a, **k: (seen.append(os.path.exists(os.path.join(d, I.FLAG))),
                                                            real(*a, **k)))
    I.write_state_guarded(d, 1000, _small_state(), None)
    assert seen == [True] and I.is_clean(d)
    assert os.path.exists(os.path.join(d, "md0000001000.r"))
    _touch(d, [I.FLAG])
    assert not I.is_clean(d)
    with pytest.raises(I.DirtyStateError):
        I.run_common(os.path.join(d, "md0000001000"), d)
    I.write_state_guarded(d, 1000, _small_state(), None)
    assert I.is_clean(d)


def test_crash_files_layout(tmp_path):
    """in.mde.after_crash as SimLoop::saveToMDE writes it (SimLoop.cxx:1175-
    1192): the atom count, per atom the 13 lines of operator<<(Atom)
    (Atom.cxx:223-240) and a blank line, the three times, the box and the
    legacy bath; mde_state.after_crash as SimLoop::saveToStream (:603-629) at
    precision 10 with the iteration after dt."""
    st = _small_state()
    I.write_after_crash(str(tmp_path), st, 1234)
    L = open(tmp_path / "in.mde.after_crash").read().split("\n")
    assert L[0] == "4"
    a0 = L[1:15]
    assert a0[0] == "12" and float(a0[1]) == pytest.approx(6 * 4.8e-10) and float(a0[2]) == pytest.approx(12 * I.AMU)
    assert a0[3] == "0 1e-08 2e-08 " and a0[4] == "0 0 0 " and a0[10] == "1" and a0[13] == ""
    assert a0[11] == f"1e-07 1e-07 {ST.NO_PBC:g} "
    last = L[1 + 3 * 14:1 + 4 * 14]
    assert last[0] == "40" and last[12] == str(1 << 2)  # the projectile tag bit
    tail = L[1 + 4 * 14:]
    assert [float(q) for q in tail[:3]] == [3e-14, 1e-12, 1e-12]
    assert tail[3] == f"1e-07 1e-07 {ST.NO_PBC:g} " and tail[4:9] == ["0.01", "0", "-5e-08", "0", "1e+13"]
    S_ = open(tmp_path / "mde_state.after_crash").read().split("\n")
    assert S_[0] == "4" and S_[1] == "12"
    t = S_[1 + 4 * 14 + 1 + 7:]
    assert float(t[0]) == 3e-14 and float(t[4]) == 2e-16 and t[5] == "1234" and t[6] == "1000"
