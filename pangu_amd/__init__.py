"""pangu_amd: MI355X (gfx950) FP64 GRMHD per-RK-stage update path.

The compute path is the HIP library ``libpangu_hip.so`` built in-tree by
``pangu_amd.build``; this package only binds it.  There is no CPU fallback:
loading fails loudly when the library is missing.
"""
from . import configs  # noqa: F401

__all__ = ["configs", "lib_path"]


def lib_path():
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpangu_hip.so")
