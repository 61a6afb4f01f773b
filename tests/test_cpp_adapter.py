"""The C++ host adapter (include/mdtk_gpu.hpp) driven like mdtk's SimLoop.

CPU: the adapter compiles against the headers and links libmdg.so.
GPU: a host-driven Verlet run through the adapter (positions and
incDisplacement pushed every step) reproduces the oracle's trajectory.
"""
import os
import subprocess

import numpy as np
import pytest

from mdtk_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_check.cpp")


def build_adapter(out_dir):
    exe = os.path.join(out_dir, "adapter_check")
    libdir = os.path.join(ROOT, "mdtk_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), SRC,
           "-L" + libdir, "-lmdg", "-Wl,-rpath," + libdir, "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_adapter_compiles_and_links(tmp_path):
    exe = build_adapter(str(tmp_path))
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_adapter_host_driven_verlet_matches_oracle(tmp_path):
    from oracle.oracle import Oracle, POT_REBO, POT_AIREBO, POT_ETORS
    exe = build_adapter(str(tmp_path))
    s = S.jitter(S.polyethylene(2, 3, 5), 0.03, 9)
    s.vel = S.maxwell_velocities(s, 300.0, stream=4)
    e, x, v, p, fx = s.to_cgs()
    mass = np.array([S.MASS_AMU[int(q)] for q in e]) * S.AMU
    inp = tmp_path / "sys.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(s.n).tobytes())
        f.write(e.astype(np.int32).tobytes())
        f.write(x.tobytes())
        f.write(v.tobytes())
        f.write(p.tobytes())
        f.write(mass.tobytes())
    out = tmp_path / "out.bin"
    steps, dt = 20, 0.2e-15
    r = subprocess.run([exe, str(inp), str(steps), repr(dt), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = out.read_bytes()
    ne = int(np.frombuffer(raw[:4], np.int32)[0])
    en = np.frombuffer(raw[4:4 + 8 * ne], np.float64)
    xf = np.frombuffer(raw[4 + 8 * ne:], np.float64).reshape(-1, 3)
    o = Oracle(e, x, v, p, None, pots=POT_REBO | POT_AIREBO | POT_ETORS)
    ro = o.run(steps, fixed_dt=dt)
    xo = o.state()[0]
    assert ne == steps + 1
    assert np.abs(en[1:] - ro["epot"]).max() <= 1e-10 * np.abs(ro["epot"]).max()
    assert np.abs(xf - xo).max() <= 1e-9 * S.ANG
