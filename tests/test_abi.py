"""The C-ABI library loads and exports every entry point include/mdg.h declares
(CPU-only: no compute call is made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        out |= set(re.findall(r"^\s*(?:int|void|const char\*)\s+\**(mdg_\w+)\s*\(", txt, re.M))
    return out


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("mdg_create", "mdg_destroy", "mdg_set_system", "mdg_energy_grad", "mdg_step_nve",
              "mdg_nl_export", "mdg_last_error"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    lib = os.path.join(ROOT, "mdtk_amd", "libmdg.so")
    if not os.path.exists(lib):
        from mdtk_amd import build
        build.build()
    L = ctypes.CDLL(lib)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(L, s)]
    assert not missing, missing


def test_engine_refuses_without_library(tmp_path):
    from mdtk_amd import engine
    with pytest.raises(engine.EngineError):
        engine._lib = None
        engine.load_library(str(tmp_path / "nope.so"))
    engine._lib = None
