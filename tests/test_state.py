"""State files in the reference's SimLoopSaver layout (CPU).

The reference cannot be built here (its SimLoop needs GSL), so no state file
written by it is available: the layout is pinned by the record sizes and
orders SimLoopSaver::write/load use (SimLoopSaver.cxx:117-625: one raw
little-endian array per attribute, the .check record 2 x uint8 + 7 x
float64 = 58 bytes, .time 4 x float64, the sphere bath 5 x float64) and by
round trips through each compression yaatk's stream may choose.  Parity with
files the reference writes is unpinned beyond that.
"""
import gzip
import lzma
import os

import numpy as np
import pytest

from mdtk_amd import state as ST
from mdtk_amd import systems as S


def _state(n=37, bath=2):
    rng = np.random.default_rng(5)
    tags = {t: rng.random(n) < 0.3 for t in ST.TAGS}
    b = None
    if bath:
        b = dict(geometry=bath, To=300.0, gamma=1.0e13, z_min=12e-8, d_boundary=2e-8, z_min_free=-5e-8,
                 center=(1e-8, 2e-8, 3e-8), radius=9e-8)
    return ST.State(elem=rng.choice([1, 12, 40], n).astype(np.int32), pos=rng.random((n, 3)) * 2e-7,
                    vel=rng.normal(size=(n, 3)) * 1e5, acc=rng.normal(size=(n, 3)) * 1e20,
                    pbc=np.array([2.5e-7, 2.5e-7, ST.NO_PBC]), pbc_count=rng.integers(-3, 3, (n, 3)).astype(np.int32),
                    pbc_enabled=np.tile([True, True, False], (n, 1)), indices=np.arange(n, dtype=np.uint32)[::-1].copy(),
                    tags=tags, bath=b, time=dict(sim_time=1.5e-13, sim_time_final=2e-12, dt=1e-16, dt_prev=9e-17),
                    check=dict(check_force=True, net_force=(1.0, -2.0, 3.0), check_energy=False,
                               initial_energy=-1e-9, current_energy=-1.1e-9, energy_from_bath=2e-12,
                               temperature=301.5))


@pytest.mark.parametrize("compress", ["xz", "gz", None])
@pytest.mark.parametrize("bath", [0, 1, 2, 3])
def test_round_trip(tmp_path, compress, bath):
    st = _state(bath=bath)
    p = str(tmp_path / "base")
    ST.write_state(p, st, compress=compress)
    r = ST.read_state(p)
    for f in ("elem", "pos", "vel", "acc", "pbc", "pbc_count", "pbc_enabled", "indices"):
        assert np.array_equal(getattr(r, f), getattr(st, f)), f
    for t in ST.TAGS:
        assert np.array_equal(r.tags[t], st.tags[t]), t
    assert r.time == st.time
    assert r.check["net_force"] == st.check["net_force"] and r.check["temperature"] == st.check["temperature"]
    if bath:
        for k in ("geometry", "To", "gamma"):
            assert r.bath[k] == st.bath[k]
        if bath == 2:
            assert (r.bath["z_min"], r.bath["d_boundary"], r.bath["z_min_free"]) == (12e-8, 2e-8, -5e-8)
        if bath == 3:
            assert r.bath["center"] == st.bath["center"] and r.bath["radius"] == 9e-8
    else:
        assert r.bath is None


def test_layout_of_each_file(tmp_path):
    st = _state(n=11, bath=3)
    p = str(tmp_path / "md")
    ST.write_state(p, st, compress="xz")

    def raw(ext):
        with lzma.open(p + ext + ".xz", "rb") as f:
            return f.read()

    n = st.n
    assert raw(".z") == st.elem.astype(np.uint8).tobytes()
    assert raw(".r") == st.pos.astype("<f8").tobytes()  # x, y, z per atom
    assert len(raw(".v")) == len(raw(".a")) == 24 * n
    assert raw(".pbc_rect") == st.pbc.astype("<f8").tobytes()
    assert len(raw(".pbc_rect.count")) == 12 * n and len(raw(".pbc_rect.enabled")) == 3 * n
    assert len(raw(".indices")) == 4 * n
    for t in ST.TAGS:
        assert len(raw(".tag." + t)) == n
    assert len(raw(".thermal_bath.common")) == 16
    assert len(raw(".thermal_bath.sphere")) == 40
    assert not os.path.exists(p + ".thermal_bath.box.xz")
    assert len(raw(".time")) == 32
    c = raw(".check")
    assert len(c) == 58 and c[0] == 1 and c[25] == 0


def test_missing_optional_files_and_bad_counts(tmp_path):
    p = str(tmp_path / "min")
    n = 5
    with gzip.open(p + ".z.gz", "wb") as f:
        f.write(bytes([12] * n))
    with open(p + ".r", "wb") as f:
        f.write(np.zeros((n, 3)).tobytes())
    st = ST.read_state(p)
    assert st.n == n and st.bath is None and st.tags["thermal_bath_applicable"].all() and not st.tags["fixed"].any()
    assert np.array_equal(st.indices, np.arange(n))
    with open(p + ".v", "wb") as f:
        f.write(np.zeros((n + 1, 3)).tobytes())
    with pytest.raises(ValueError):
        ST.read_state(p)


def test_state_of_a_system_round_trips(tmp_path):
    """A slab built here, written in the layout and read back, gives the same
    CGS arrays the engine would be handed directly."""
    s = S.pe_slab(1)
    s.vel = S.maxwell_velocities(s, 300.0, stream=2)
    e, x, v, pbc, fx = s.to_cgs()
    n = s.n
    st = ST.State(e, x.reshape(n, 3), v.reshape(n, 3), np.zeros((n, 3)), pbc, np.zeros((n, 3), np.int32),
                  np.tile(pbc != ST.NO_PBC, (n, 1)), np.arange(n, dtype=np.uint32),
                  {"fixed": fx.astype(bool)})
    p = str(tmp_path / "slab")
    ST.write_state(p, st)
    r = ST.read_state(p)
    assert np.array_equal(r.pos, x.reshape(n, 3)) and np.array_equal(r.vel, v.reshape(n, 3))
    assert np.array_equal(r.tags["fixed"], fx.astype(bool))
