"""Trajectory files (CPU): the SnapshotList state file and XVA text frames
written from captured frames (SnapshotList.hpp:38-204, SimLoop.cxx:681-748)."""
import numpy as np

from mdtk_amd import trajectory as T
from mdtk_amd.engine import FRAME_DTYPE, RECORD_DTYPE, FRAME_SNAPSHOT, FRAME_XVA


def _frames():
    rng = np.random.default_rng(2)
    rec = np.zeros(9, RECORD_DTYPE)
    rec["x"] = rng.random((9, 3)) * 3e-7
    rec["v"] = rng.normal(size=(9, 3)) * 1e5
    rec["pbc_count"] = rng.integers(-2, 3, (9, 3))
    rec["atom"] = [7, 9, 7, 9, 0, 1, 2, 3, 4]
    fr = np.zeros(4, FRAME_DTYPE)
    fr[0] = (FRAME_SNAPSHOT, 0, 0, 0.0, 1e-20, 0, 2)
    fr[1] = (FRAME_SNAPSHOT, 0, 31, 1.0e-14, 3e-16, 2, 2)
    fr[2] = (FRAME_XVA, 0, 0, 0.0, 1e-20, 4, 5)
    fr[3] = (FRAME_SNAPSHOT, 1, 5, 2e-15, 3e-16, 0, 2)
    return fr, rec


def test_snapshots_conf_layout_and_round_trip(tmp_path):
    fr, rec = _frames()
    sel, shots = T.snapshot_list(fr, rec, system=0)
    assert list(sel) == [7, 9] and [t for t, _ in shots] == [0.0, 1.0e-14]
    assert np.array_equal(shots[1][1]["pos"], rec["x"][2:4].astype(np.float32))
    p = str(tmp_path / "snapshots.conf")
    T.write_snapshots_conf(p, sel, shots)
    raw = open(p, "rb").read()
    assert len(raw) == 4 + 2 * 4 + 4 + 2 * (8 + 4 + 2 * 36)
    assert T.ATOM_SNAPSHOT.itemsize == 36
    sel2, shots2 = T.read_snapshots_conf(p)
    assert np.array_equal(sel2, sel)
    for (t1, a1), (t2, a2) in zip(shots, shots2):
        assert t1 == t2 and np.array_equal(a1, a2)


def test_xva_text():
    fr, rec = _frames()
    f = T.xva_frames(fr, 0)[0]
    lines = T.xva_text(f, rec).splitlines()
    r = rec[4:9]
    mean = np.abs(r["v"]).sum() / 15
    scale = 10.0 ** (np.ceil(np.log10(mean)) - 2)
    assert float(lines[0]) == scale and lines[1] == "1e-08" and lines[2] == "5"
    assert lines[3] == " ".join(f"{q / scale:.2f}" for q in r["v"][0])
    assert lines[4] == " ".join(f"{q * 1e8:.2f}" for q in r["x"][0])
    assert lines[5] == " ".join(str(q) for q in r["pbc_count"][0])
    assert len(lines) == 3 + 3 * 5 + 7 + 4
    assert lines[-4] == f"{0.0:.10e}" and lines[-2] == "0"
    assert T.xva_name(f) == "mde0000000000.xva"


def test_record_layouts_match_the_header():
    """mdg_frame is 48 bytes and mdg_atom_record 64 (include/mdg.h); the
    Python views must read them field for field."""
    assert FRAME_DTYPE.itemsize == 48 and RECORD_DTYPE.itemsize == 64
    assert [FRAME_DTYPE.fields[k][1] for k in ("kind", "system", "iteration", "sim_time", "dt", "first", "count")] \
        == [0, 4, 8, 16, 24, 32, 40]
    assert [RECORD_DTYPE.fields[k][1] for k in ("x", "v", "pbc_count", "atom")] == [0, 24, 48, 60]
