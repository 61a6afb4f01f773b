"""MI355X-native force-and-integration engine for MDTK-style hydrocarbon MD.

The compute path is hand-written HIP for gfx950 in ``libmdg.so`` behind the C
ABI of ``include/mdg.h``; this package is the Python front end of that ABI
(ctypes), mirroring the reference's potential chain and Verlet loop.  There is
no CPU fallback: importing the engine on a machine without the built library
or without a GPU raises.
"""
from .engine import (Engine, EngineError, POT_FBZL, POT_BRENNER, POT_REBO, POT_AIREBO, POT_ETORS,
                     load_library)

__all__ = ["Engine", "EngineError", "POT_FBZL", "POT_BRENNER", "POT_REBO", "POT_AIREBO", "POT_ETORS",
           "load_library"]
