"""Parthenon-style HDF5 output of a HipSim's primitives (include/pangu_phdf.h).

Variable names follow the reference's field registry
(package_registration.cc:84-89): density, energy, weighted_velocity (3),
magnetic_field (3), entropy, electron_entropy.  The data are copied off the
GPU and written by libpangu_phdf.so, which links HDF5; it fails loudly when
that library was not built."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Mesh(C.Structure):
    _fields_ = [("nblocks", C.c_int), ("nx", C.c_int * 3), ("ng", C.c_int), ("ndim", C.c_int),
                ("ncycle", C.c_int), ("time", C.c_double), ("dt", C.c_double),
                ("root_nx", C.c_int * 3), ("root_min", C.c_double * 3), ("root_max", C.c_double * 3),
                ("lx", C.POINTER(C.c_int64)), ("xmin", C.POINTER(C.c_double)),
                ("dx", C.POINTER(C.c_double)), ("level", C.POINTER(C.c_int))]


class Var(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ncomp", C.c_int), ("data", C.POINTER(C.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libpangu_phdf.so")
        if not os.path.exists(path):
            raise RuntimeError("libpangu_phdf.so is not built (python -m pangu_amd.build; needs HDF5)")
        L = C.CDLL(path)
        L.pangu_write_phdf.argtypes = [C.c_char_p, C.POINTER(Mesh), C.c_int, C.POINTER(Var)]
        L.pangu_phdf_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def variable_slices(cfg):
    """name -> (first primitive component, component count) for a config."""
    out = {"density": (0, 1), "energy": (1, 1), "weighted_velocity": (2, 3)}
    n = 5
    if cfg.get("enable_B", 1):
        out["magnetic_field"] = (n, 3)
        n += 3
    out["entropy"] = (n, 1)
    n += 1
    if cfg.get("enable_heating", 0):
        out["electron_entropy"] = (n, 1)
    return out


def write(sim, path, variables=None, container=0, mesh=0, ncycle=0):
    """Writes mesh `mesh` of `sim` (container 0 = base) to path and
    path.xdmf.  variables: names of variable_slices (default: all)."""
    cfg = sim.cfgs[mesh]
    slices = variable_slices(cfg)
    names = list(slices) if variables is None else list(variables)
    for nm in names:
        if nm not in slices:
            raise ValueError(f"unknown variable {nm!r} (have {sorted(slices)})")
    dt, tm, first, nb = sim.mesh_state(mesh)
    flat = sim.get("prim", container)
    per_block = sim.blocks(flat.reshape(-1)) if sim.uniform else sim.blocks(flat)
    lx = np.zeros((nb, 3), dtype=np.int64)
    level = np.zeros(nb, dtype=np.int32)
    xmin = np.zeros((nb, 3))
    dx = np.zeros((nb, 3))
    cells = []
    for bi in range(nb):
        b = first + bi
        xd, n = sim.block_geom(b)
        s = n[3:6]
        loc = sim.block_loc(b)
        level[bi] = loc[0]
        for d in range(3):
            dx[bi, d] = xd[3 + d]
            xmin[bi, d] = xd[d] + s[d] * xd[3 + d]
            lx[bi, d] = loc[1 + d]
        cells.append(sim.block_interior(per_block[b], b))
    nxb = [int(cells[0].shape[3]), int(cells[0].shape[2]), int(cells[0].shape[1])]
    arrays, vs = [], (Var * len(names))()
    for v, nm in enumerate(names):
        q0, nc = slices[nm]
        a = np.ascontiguousarray(np.stack([c[q0:q0 + nc] for c in cells]))
        if nc == 1:
            a = np.ascontiguousarray(a[:, 0])
        arrays.append(a)
        vs[v] = Var(nm.encode(), nc, a.ctypes.data_as(C.POINTER(C.c_double)))
    m = Mesh()
    m.nblocks = nb
    for d in range(3):
        m.nx[d] = nxb[d]
        m.root_nx[d] = cfg["nx"][d]
        m.root_min[d] = cfg["xmin"][d]
        m.root_max[d] = cfg["xmax"][d]
    m.ng = cfg.get("ng", 4)
    m.ndim = sum(1 for d in range(3) if cfg["nx"][d] > 1)
    m.ncycle = ncycle
    m.time = tm
    m.dt = dt
    m.lx = lx.ctypes.data_as(C.POINTER(C.c_int64))
    m.xmin = xmin.ctypes.data_as(C.POINTER(C.c_double))
    m.dx = dx.ctypes.data_as(C.POINTER(C.c_double))
    m.level = level.ctypes.data_as(C.POINTER(C.c_int))
    if lib().pangu_write_phdf(os.fsencode(path), C.byref(m), len(names), vs) != 0:
        raise RuntimeError(lib().pangu_phdf_last_error().decode())
    return path
