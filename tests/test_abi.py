"""The C-ABI library on the CPU: it loads, exports every declared symbol, its
ctypes mirror matches the C layout, configuration errors are rejected
before any device call, and the host-built geometry matches the oracle."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_py as O
from pangu_amd import abi, configs, sim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("pangu_hip.h", "pangu_sim.h", "pangu_mlp.h")]


def _declared(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(pangu_\w+)\s*\(", src, flags=re.M)))


@pytest.fixture(scope="module")
def L():
    return abi.lib()


def test_every_declared_function_is_exported(L):
    names = sum((_declared(h) for h in HEADERS), [])
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_ctypes_mirror_covers_task_header():
    assert sorted(abi.PROTOTYPES) == _declared(HEADERS[0])


def test_struct_layouts_match_c(tmp_path):
    """sizeof/offsetof of every ABI struct, compiled from the header by gcc,
    equal the ctypes mirrors used by the Python host side."""
    checks = {
        "pangu_config": (abi.PanguConfig, ["metric", "gamma", "heat_model", "pexcise", "nb", "bc"]),
        "pangu_block": (abi.PanguBlock, ["nx", "ng", "lx", "xmin", "dx", "level"]),
        "pangu_container": (abi.PanguContainer, ["prim", "cons", "flux"]),
        "pangu_shared": (abi.PanguShared, ["alfven", "mask", "flag_recovery"]),
        "pangu_sim_cfg": (sim.SimCfg, ["nx", "bc", "fel_0", "cache_geometry", "seed", "nthreads"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pangu_hip.h"',
             '#include "pangu_sim.h"', "int main(void) {"]
    for st, (_, fields) in checks.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fields:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for st, (cls, fields) in checks.items():
        assert int(got[st]) == C.sizeof(cls), st
        for f in fields:
            assert int(got[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_sim_cfg_mirrors_oracle_cfg():
    assert [n for n, _ in sim.SimCfg._fields_] == [n for n, _ in O.OracleCfg._fields_]


def _create(L, d, blocks=None, nblocks=None):
    cfg = abi.config_from(d)
    blk = abi.blocks_from(d) if blocks is None else blocks
    n = len(blk) if nblocks is None else nblocks
    ctx = C.c_void_p()
    return L.pangu_create(C.byref(cfg), blk, n, 0, C.byref(ctx))


@pytest.mark.parametrize("field,value", [
    ("riemann", 2), ("limiter", 5), ("recovery", 3), ("heat_model", 6), ("gamma", 1.0),
    ("metric", 4),
])
def test_create_rejects_invalid_configuration(L, field, value):
    d = configs.torus2d(64, 32)
    d[field] = value
    assert _create(L, d) == abi.EINVAL


def test_create_rejects_bad_blocks(L):
    d = configs.torus2d(64, 32)
    d["nb"] = (2, 2, 1)
    assert _create(L, d, nblocks=0) == abi.EINVAL
    blk = abi.blocks_from(d)
    blk[1].nx[0] += 1  # blocks of one context share a shape
    assert _create(L, d, blocks=blk) == abi.EINVAL
    blk = abi.blocks_from(d)
    for b in blk:
        b.ng[0] = 2  # ppm4 reads three cells on each side
    assert _create(L, d, blocks=blk) == abi.EINVAL


POINTS = {
    "minkowski": [(0.0, 0.3, 0.7, 0.0)],
    "bl": [(0.0, 4.0, 1.1, 0.3), (0.0, 12.0, 0.4, -2.0)],
    "cks": [(0.0, 3.0, -2.0, 1.5), (0.0, -6.0, 0.5, -0.25)],
    "mks": [(0.0, 0.4, 0.3, 0.0), (0.0, 2.1, -0.7, 1.0), (0.0, 1.0, 0.999, 0.0)],
}


@pytest.mark.parametrize("kind", list(POINTS))
def test_geometry_entry_matches_oracle(L, kind):
    d = configs.torus2d(64, 32)
    d["metric"] = abi.METRIC[kind]
    cfg = abi.config_from(d)
    for x in POINTS[kind]:
        xa = np.array(x, dtype=np.float64)
        cen = np.zeros(abi.CENTRE_NCOMP)
        assert L.pangu_geometry_entry(C.byref(cfg), 0, xa.ctypes.data_as(C.POINTER(C.c_double)),
                                      cen.ctypes.data_as(C.POINTER(C.c_double))) == abi.OK
        gcov, gcon, gdet, ok = O.metric_at(d["metric"], xa, d["a"], d["h"])
        conn = O.christoffel(d["metric"], xa, d["a"], d["h"])
        assert ok
        np.testing.assert_array_equal(cen[0:16], gcov.ravel())
        np.testing.assert_array_equal(cen[16:32], gcon.ravel())
        assert cen[32] == np.sqrt(abs(gdet))
        assert cen[33] == 1.0 / np.sqrt(-gcon[0, 0])
        assert cen[34] == np.sqrt(-1.0 / gcon[0, 0])
        np.testing.assert_array_equal(cen[38:102], conn.ravel())
        for dd in range(3):
            face = np.zeros(abi.FACE_NCOMP)
            assert L.pangu_geometry_entry(C.byref(cfg), dd + 1,
                                          xa.ctypes.data_as(C.POINTER(C.c_double)),
                                          face.ctypes.data_as(C.POINTER(C.c_double))) == abi.OK
            np.testing.assert_array_equal(face[0:16], gcov.ravel())
            np.testing.assert_array_equal(face[16:20], gcon[0])
            assert face[20] == gcon[dd + 1, dd + 1]
            assert face[23] == 1.0


def test_geometry_entry_floors(L):
    d = configs.torus2d(64, 32)
    cfg = abi.config_from(d)
    for x1 in (-0.5, 0.3, 1.7):
        xa = np.array([0.0, x1, 0.2, 0.0])
        cen = np.zeros(abi.CENTRE_NCOMP)
        L.pangu_geometry_entry(C.byref(cfg), 0, xa.ctypes.data_as(C.POINTER(C.c_double)),
                               cen.ctypes.data_as(C.POINTER(C.c_double)))
        r = np.exp(x1)
        assert cen[37] == r
        assert cen[35] == d["density_floor"] * max(r, 1.0) ** d["density_floor_pow"]
        assert cen[36] == d["energy_floor"] * max(r, 1.0) ** d["energy_floor_pow"]


def test_geometry_entry_polar_face_is_invalid(L):
    """MKS at theta = 0 exactly: the metric is singular, the face is marked
    invalid and its lapse is zero (the flux there is set to zero)."""
    d = configs.torus2d(64, 32)
    cfg = abi.config_from(d)
    xa = np.array([0.0, 1.0, -1.0, 0.0])
    face = np.zeros(abi.FACE_NCOMP)
    L.pangu_geometry_entry(C.byref(cfg), 2, xa.ctypes.data_as(C.POINTER(C.c_double)),
                           face.ctypes.data_as(C.POINTER(C.c_double)))
    assert face[23] == 0.0 and face[22] == 0.0
    assert O.metric_at(d["metric"], xa, d["a"], d["h"])[3] == 0


def test_sim_binding_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(sim, "_lib", None)
    monkeypatch.setattr(sim, "lib_path", lambda: str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError):
        sim.lib()


def _refined_blocks(d, leaves):
    """pangu_block list of (level, lx1, lx2) leaves of a 2-D configuration."""
    out = (abi.PanguBlock * len(leaves))()
    nbk = [d["nx"][i] // d["nb"][i] for i in range(3)]
    for b, (level, i, j) in enumerate(leaves):
        B = out[b]
        B.level = level
        for dd, l in enumerate((i, j, 0)):
            cnt = d["nb"][dd] << level if d["nx"][dd] > 1 else 1
            x0, x1 = d["xmin"][dd], d["xmax"][dd]
            B.nx[dd] = nbk[dd]
            B.ng[dd] = d["ng"] if d["nx"][dd] > 1 else 0
            B.lx[dd] = l
            B.xmin[dd] = x0 + (x1 - x0) * l / cnt
            B.dx[dd] = (x1 - x0) / cnt / nbk[dd]
    return out


def _quadrant_refined(child_level=1):
    roots = [(0, 0, 0), (0, 1, 0), (0, 0, 1)]
    kids = [(child_level, i, j) for j in (2, 3) for i in (2, 3)]
    return roots, kids


def test_create_checks_refined_meshes(L):
    """Refined meshes are checked on the host before any device work:
    blocks must tile the domain, neighbours may differ by one level, and a
    block next to a coarser one may not touch an outflow face (its
    prolongation stencil would need cells beyond the boundary).  A valid
    mesh passes the checks (it then fails only for want of a GPU here)."""
    d = configs.orszag_tang(32, (2, 2, 1))
    roots, kids = _quadrant_refined()

    def create(leaves, dd=d):
        cfg = abi.config_from(dd)
        blk = _refined_blocks(dd, leaves)
        ctx = C.c_void_p()
        r = L.pangu_create(C.byref(cfg), blk, len(leaves), 0, C.byref(ctx))
        if r == abi.OK:
            L.pangu_destroy(ctx)
        return r

    assert create(roots + kids) != abi.EINVAL                      # periodic box: valid
    assert create(roots + kids[:3]) == abi.EINVAL                  # a gap
    assert create(roots + [(0, 1, 1)] + kids) == abi.EINVAL        # a block inside another
    assert create(roots + [(1, 4, 4)] + kids) == abi.EINVAL        # outside the level-1 grid
    deep = roots + kids[1:] + [(2, i, j) for j in (4, 5) for i in (4, 5)]
    assert create(deep) == abi.EINVAL                              # level 2 next to level 0
    out = dict(d, bc=[[configs.OUTFLOW] * 2, [configs.OUTFLOW] * 2, [configs.PERIODIC] * 2])
    assert create(roots + kids, out) == abi.EINVAL                 # fine block on an outflow face
