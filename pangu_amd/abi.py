"""ctypes mirror of the task-level C ABI (include/pangu_hip.h).

The structures and prototypes here match the header one to one, so a
Python caller can drive the same entry points a Parthenon task list binds.
Loading raises when libpangu_hip.so has not been built.
"""
import ctypes as C
import os

from . import lib_path


class PanguConfig(C.Structure):
    _fields_ = [
        ("metric", C.c_int), ("a", C.c_double), ("h", C.c_double),
        ("gamma", C.c_double), ("cfl", C.c_double),
        ("riemann", C.c_int), ("limiter", C.c_int), ("recovery", C.c_int),
        ("enable_B", C.c_int), ("enable_heating", C.c_int),
        ("density_floor", C.c_double), ("energy_floor", C.c_double),
        ("density_floor_pow", C.c_double), ("energy_floor_pow", C.c_double),
        ("sigma_max", C.c_double), ("lorentz_max", C.c_double),
        ("heat_model", C.c_int),
        ("fel_constant", C.c_double), ("gamma_e", C.c_double), ("gamma_p", C.c_double),
        ("ratio_min", C.c_double), ("ratio_max", C.c_double),
        ("limit_kel", C.c_int), ("suppress_highb_heat", C.c_int),
        ("enforce_positive_dissipation", C.c_int),
        ("excise", C.c_int), ("r_excise", C.c_double), ("dexcise", C.c_double),
        ("pexcise", C.c_double),
        ("nb", C.c_int * 3), ("bc", (C.c_int * 2) * 3),
    ]


class PanguBlock(C.Structure):
    _fields_ = [("nx", C.c_int * 3), ("ng", C.c_int * 3), ("lx", C.c_int * 3),
                ("xmin", C.c_double * 3), ("dx", C.c_double * 3), ("level", C.c_int)]


_PPD = C.POINTER(C.POINTER(C.c_double))


class PanguContainer(C.Structure):
    _fields_ = [("prim", _PPD), ("cons", _PPD), ("flux", _PPD * 3)]


class PanguShared(C.Structure):
    _fields_ = [("alfven", _PPD), ("emf", _PPD), ("flag", _PPD),
                ("mask", C.POINTER(C.POINTER(C.c_uint8))),
                ("c2p_status", C.POINTER(C.POINTER(C.c_int32))),
                ("c2p_iters", C.POINTER(C.POINTER(C.c_int32))),
                ("flag_recovery", _PPD), ("c2p_root", _PPD)]


# Error codes and enumerations of the header.
OK, EINVAL, EDEVICE, ENOMEM, ESTATE = 0, -1, -2, -3, -4
METRIC = dict(minkowski=0, bl=1, cks=2, mks=3)
RIEMANN = dict(laxf=0, hll=1)
LIMITER = dict(mc=0, ppm4=1)
RECOVERY = dict(chainer=0, robuster=1, onedvsq=2)
HEAT = dict(constant=0, howes=1, kawazura=2, werner=3, rowan=4, sharma=5)
BC = dict(periodic=0, outflow=1, reflect=2)

# Table component counts of pangu_geometry_entry (centre, face).
CENTRE_NCOMP = 102
FACE_NCOMP = 24

_vp, _dp, _ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
PROTOTYPES = {
    "pangu_create": (C.c_int, [C.POINTER(PanguConfig), C.POINTER(PanguBlock), C.c_int, C.c_int,
                               C.POINTER(_vp)]),
    "pangu_create_batch": (C.c_int, [C.POINTER(PanguConfig), C.c_int, C.POINTER(PanguBlock),
                                     C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(_vp)]),
    "pangu_destroy": (C.c_int, [_vp]),
    "pangu_last_error": (C.c_char_p, [_vp]),
    "pangu_geometry_entry": (C.c_int, [C.POINTER(PanguConfig), C.c_int, _dp, _dp]),
    "pangu_register_container": (C.c_int, [_vp, C.POINTER(PanguContainer), _ip]),
    "pangu_register_shared": (C.c_int, [_vp, C.POINTER(PanguShared)]),
    "pangu_fix_primitive": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_calculate_conservative": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_calculate_fluxes": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_constrained_transport": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_flux_divergence": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    "pangu_update_independent": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, _vp]),
    "pangu_add_geometric_source": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    "pangu_recovery": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_apply_electron_heating": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_fix_recovery": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_estimate_timestep": (C.c_int, [_vp, C.c_int, _dp, _vp]),
    "pangu_exchange_ghosts": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_flux_correction": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_apply_boundary_conditions": (C.c_int, [_vp, C.c_int, _vp]),
    "pangu_set_tile_order": (C.c_int, [_vp, C.c_int, C.c_int]),
    "pangu_get_tile_order": (C.c_int, [_vp, C.c_int]),
    "pangu_set_split_source": (C.c_int, [_vp, C.c_int]),
    "pangu_set_share_reconstruction": (C.c_int, [_vp, C.c_int]),
    "pangu_set_pair_flux": (C.c_int, [_vp, C.c_int]),
    "pangu_set_flux_march": (C.c_int, [_vp, C.c_int, C.c_int]),
    "pangu_set_tile_size": (C.c_int, [_vp, C.c_int, C.c_int]),
    "pangu_set_persistent_recovery": (C.c_int, [_vp, C.c_int]),
    "pangu_get_tile_size": (C.c_int, [_vp, C.c_int]),
    "pangu_set_fuse_ct": (C.c_int, [_vp, C.c_int]),
    "pangu_set_overlap_exchange": (C.c_int, [_vp, C.c_int]),
    "pangu_set_fuse_update_recovery": (C.c_int, [_vp, C.c_int]),
    "pangu_set_fuse_fix_x1": (C.c_int, [_vp, C.c_int]),
    "pangu_diagnostics": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), _dp, _vp]),
    "pangu_diagnostics_noalloc": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), _vp, _vp]),
    "pangu_stage": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_double,
                              C.c_int, C.c_int, _vp]),
}

_lib = None


def lib():
    """Load libpangu_hip.so with typed prototypes (raises if missing)."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python -m pangu_amd.build` first")
        L = C.CDLL(path)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def config_from(d):
    """pangu_config from a run-configuration dict (pangu_amd.configs)."""
    c = PanguConfig()
    for name, _ in PanguConfig._fields_:
        if name not in d:
            continue
        v = d[name]
        if name == "bc":
            for i in range(3):
                for s in range(2):
                    c.bc[i][s] = int(v[i][s])
        elif name == "nb":
            for i in range(3):
                c.nb[i] = int(v[i])
        else:
            setattr(c, name, v)
    return c


def blocks_from(d):
    """The meshblocks of a run configuration, in Parthenon's block order
    (x1 fastest), with the same interval arithmetic as the host driver."""
    nx, nb, ng = d["nx"], d["nb"], d["ng"]
    n = [nx[i] // nb[i] for i in range(3)]
    out = (PanguBlock * (nb[0] * nb[1] * nb[2]))()
    b = 0
    for k in range(nb[2]):
        for j in range(nb[1]):
            for i in range(nb[0]):
                B = out[b]
                for dd, l in enumerate((i, j, k)):
                    x0, x1 = d["xmin"][dd], d["xmax"][dd]
                    lo = x0 if l == 0 else _mesh_gen(l, nb[dd], x0, x1)
                    hi = x1 if l == nb[dd] - 1 else _mesh_gen(l + 1, nb[dd], x0, x1)
                    B.nx[dd] = n[dd]
                    B.ng[dd] = ng if nx[dd] > 1 else 0
                    B.lx[dd] = l
                    B.xmin[dd] = lo
                    B.dx[dd] = (hi - lo) / n[dd]
                b += 1
    return out


def _mesh_gen(idx, n, xmin, xmax):
    """Block edge idx of n on [xmin, xmax] in the symmetric form of
    Parthenon's UniformMeshGeneratorX (exact at both ends and the middle)."""
    x = ((idx - n // 2) + (idx - (n + 1) // 2)) / (2.0 * n)
    return 0.5 * (xmin + xmax) + (x * xmax - x * xmin)
