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
