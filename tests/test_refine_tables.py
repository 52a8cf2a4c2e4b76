"""The refinement gather tables of the product (refine_host.cpp), checked on
the CPU against the oracle: every index the GPU kernels would touch is
inside its block, and the kernels' arithmetic replayed over the tables
reproduces the oracle's cross-level exchange and flux correction bit for bit
(tests/refine_tables_check.cpp).  No GPU is needed: this is the check that
keeps an out-of-bounds table entry from ever reaching a kernel."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
from pangu_amd import configs
from test_refinement import cks_smr_mesh
from test_refinement_gpu import CASES

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("refine_check") / "librefine_check.so")
    csrc = os.path.join(_ROOT, "pangu_amd", "csrc")
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           f"-I{csrc}", os.path.join(_ROOT, "tests", "refine_tables_check.cpp"),
           os.path.join(csrc, "refine_host.cpp"), os.path.join(csrc, "problems.cpp"),
           os.path.join(csrc, "geometry.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    L = C.CDLL(out)
    L.check_tables.restype = C.c_int
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


CHECK_CASES = dict(CASES)
CHECK_CASES["fm_torus_cks_tree"] = dict(cks_smr_mesh(), metric=0, excise=0)
# a refined quadrant at the corner of a periodic box: stencils wrap
CHECK_CASES["ot2d_periodic_corner"] = dict(configs.orszag_tang(32, (2, 2, 1)),
                                           refine=[(1, (0.5, 0.5, -0.5), (1.0, 1.0, 0.5))])


@pytest.mark.parametrize("name", sorted(CHECK_CASES))
def test_tables_replay_the_oracle(checker, name):
    cfg = CHECK_CASES[name]
    o = O.OracleSim(cfg, nthreads=4)
    o.init()
    nb = o.nblocks
    ib = np.zeros((nb, 10), dtype=np.int32)
    db = np.zeros((nb, 6))
    for b in range(nb):
        xd, n = o.block_geom(b)
        loc = o.block_loc(b)
        for d in range(3):
            ib[b, d] = n[6 + d] - n[3 + d] + 1
            ib[b, 3 + d] = n[3 + d]
            ib[b, 6 + d] = loc[1 + d]
            db[b, d] = xd[d] + n[3 + d] * xd[3 + d]
            db[b, 3 + d] = xd[3 + d]
        ib[b, 9] = loc[0]
    # random primitives everywhere, so every ghost the exchange writes changes
    P = o.get("prim", 1)
    P[...] = np.random.default_rng(1).standard_normal(P.shape)
    o.set("prim", P, 1)
    before = np.ascontiguousarray(o.get("prim", 1))
    o.task("exchange", 0, 1)
    after = np.ascontiguousarray(o.get("prim", 1))
    assert (before != after).any()
    for t in ("fix_primitive", "conservative", "fluxes"):
        o.task(t, 0)
    if cfg.get("enable_B", 1):
        o.task("ct", 0)
    fb = [np.ascontiguousarray(o.get(f"flux{d + 1}")) for d in range(3)]
    o.task("flux_correction", 0, 0)
    fa = [np.ascontiguousarray(o.get(f"flux{d + 1}")) for d in range(3)]
    nbv = np.array(cfg["nb"], dtype=np.int32)
    bcv = np.array(cfg["bc"], dtype=np.int32).reshape(-1)
    counts = np.zeros(3, dtype=np.int64)
    r = checker.check_tables(nb, _ip(ib), _dp(db), _ip(nbv), _ip(bcv), P.shape[1], _dp(before),
                             _dp(after), fb[0].shape[1], _dp(fb[0]), _dp(fb[1]), _dp(fb[2]),
                             _dp(fa[0]), _dp(fa[1]), _dp(fa[2]),
                             counts.ctypes.data_as(C.POINTER(C.c_long)))
    print(name, "blocks", nb, "fill/prolong/flux entries", counts.tolist())
    assert r == 0
    assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0
