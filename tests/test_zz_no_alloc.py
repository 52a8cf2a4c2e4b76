"""pangu_diagnostics_noalloc (the allocation-free variant of
pangu_diagnostics) against pangu_diagnostics, through the C ABI (GPU).

Two calls in a row on different contexts, the second on a smaller mesh, so
the shared static workspace holds the first call's partials when the second
starts; each must equal the allocating entry point bit for bit.  A context
whose interior needs more workspace than the static 8 MiB is rejected with
an error before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from pangu_amd import abi, configs
from pangu_amd.sim import DIAG_NAMES, HipSim, lib

ND = len(DIAG_NAMES)


class _Hip:
    """The HIP runtime libpangu_hip.so links (loaded with it): device
    buffers for the entry point's output."""

    def __init__(self):
        lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def to_host(self, p, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # D2H
        return out


def _noalloc(h, hip, shell, container=0):
    L = abi.lib()  # typed prototypes of include/pangu_hip.h
    ctx, handles, stream = h.context()
    sh = np.asarray([shell] * h.nmesh if np.isscalar(shell) else shell, dtype=np.int32)
    out = hip.alloc(h.nmesh * ND * 8)
    try:
        assert hip.h.hipMemset(out, 0xFF, h.nmesh * ND * 8) == 0  # NaN bytes: every entry must be written
        r = L.pangu_diagnostics_noalloc(ctx, handles[container], sh.ctypes.data_as(C.POINTER(C.c_int)),
                                        out, stream)
        if r != 0:
            return r, L.pangu_last_error(ctx).decode(), None
        return 0, "", hip.to_host(out, h.nmesh * ND)
    finally:
        hip.h.hipFree(out)


def _reference(h, shell, container=0):
    res = h.diagnostics(shell, container)
    res = res if isinstance(res, list) else [res]
    return np.array([[m[k] for k in DIAG_NAMES] for m in res]).reshape(-1)


@pytest.mark.gpu
def test_noalloc_diagnostics_equal_allocating_entry_point_twice():
    hip = _Hip()
    runs = [
        ("torus3d_64x32x32_8blk", configs.torus3d((64, 32, 32), (2, 2, 2)), 20, 1),
        ("ot32", configs.orszag_tang(32), 5, 0),  # smaller: fewer tiles than the first call
    ]
    for name, cfg, shell, container in runs:
        print("case", name, flush=True)
        h = HipSim(cfg)
        h.init()
        h.step(2, "graph")  # returns after the stream has finished
        r, msg, got = _noalloc(h, hip, shell, container)
        assert r == 0, msg
        ref = _reference(h, shell, container)
        print(name, "noalloc", got, "\nreference", ref, flush=True)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), name
        h.close()


@pytest.mark.gpu
def test_noalloc_diagnostics_batch_matches():
    """A batched context (several meshes, per-mesh shells) after a larger call."""
    hip = _Hip()
    ens = [configs.torus2d(64, 32), configs.torus2d(32, 16), configs.torus2d(48, 32)]
    h = HipSim(ens)
    h.init()
    h.step(1, "graph")
    shells = [10, 3, -1]
    r, msg, got = _noalloc(h, hip, shells)
    assert r == 0, msg
    ref = _reference(h, shells)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    h.close()


@pytest.mark.gpu
def test_noalloc_diagnostics_rejects_a_context_beyond_the_workspace():
    """256x256x520 Minkowski zones = 133 120 interior tiles x 8 partials >
    2^20 doubles: rejected with a message, before any device work (the
    fields are allocated but never initialised)."""
    hip = _Hip()
    cfg = configs.orszag_tang(256, (4, 4, 8))
    cfg.update(nx=[256, 256, 520], xmin=[0.0, 0.0, 0.0], xmax=[1.0, 1.0, 2.03125])
    h = HipSim(cfg)
    r, msg, _ = _noalloc(h, hip, -1)
    print(r, msg, flush=True)
    assert r == -1 and "static workspace holds" in msg
    h.close()
