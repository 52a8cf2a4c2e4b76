"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/ and by bench.py's cpu_baseline leg; the product package
never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


MAX_REF = 8  # ORACLE_MAX_REF


class OracleCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int * 3), ("nb", C.c_int * 3), ("ng", C.c_int),
        ("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("bc", (C.c_int * 2) * 3),
        ("dt_init", C.c_double), ("dt_factor", C.c_double),
        ("metric", C.c_int), ("a", C.c_double), ("h", C.c_double),
        ("gamma", C.c_double), ("cfl", C.c_double),
        ("riemann", C.c_int), ("limiter", C.c_int), ("recovery", C.c_int),
        ("enable_B", C.c_int), ("enable_heating", C.c_int),
        ("density_floor", C.c_double), ("energy_floor", C.c_double),
        ("density_floor_pow", C.c_double), ("energy_floor_pow", C.c_double),
        ("sigma_max", C.c_double), ("lorentz_max", C.c_double),
        ("heat_model", C.c_int),
        ("fel_constant", C.c_double), ("gamma_e", C.c_double), ("gamma_p", C.c_double),
        ("ratio_min", C.c_double), ("ratio_max", C.c_double), ("fel_0", C.c_double),
        ("limit_kel", C.c_int), ("suppress_highb_heat", C.c_int),
        ("enforce_positive_dissipation", C.c_int),
        ("excise", C.c_int), ("r_excise", C.c_double), ("dexcise", C.c_double),
        ("pexcise", C.c_double),
        ("stage_init", C.c_int), ("exchange_after", C.c_int), ("cache_geometry", C.c_int),
        ("problem", C.c_int),
        ("rin", C.c_double), ("rmax", C.c_double), ("kappa", C.c_double),
        ("perturbation", C.c_double), ("beta", C.c_double), ("aphi_cut", C.c_double),
        ("seed", C.c_uint64), ("nthreads", C.c_int),
        ("nref", C.c_int), ("ref_level", C.c_int * MAX_REF),
        ("ref_min", (C.c_double * 3) * MAX_REF), ("ref_max", (C.c_double * 3) * MAX_REF),
        ("magnetic_mad", C.c_int),
    ]


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.oracle_sim_create.restype = vp
        L.oracle_sim_create.argtypes = [C.POINTER(OracleCfg)]
        L.oracle_sim_destroy.argtypes = [vp]
        L.oracle_sim_init.argtypes = [vp]
        L.oracle_sim_step.argtypes = [vp, C.c_int]
        L.oracle_sim_task.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, dp]
        L.oracle_sim_field_size.restype = C.c_int64
        L.oracle_sim_field_size.argtypes = [vp, C.c_int]
        L.oracle_sim_get.argtypes = [vp, C.c_int, C.c_int, vp]
        L.oracle_sim_set.argtypes = [vp, C.c_int, C.c_int, vp]
        L.oracle_sim_time.restype = C.c_double
        L.oracle_sim_time.argtypes = [vp]
        L.oracle_sim_dt.restype = C.c_double
        L.oracle_sim_dt.argtypes = [vp]
        L.oracle_sim_ncycle.argtypes = [vp]
        L.oracle_sim_block_geom.argtypes = [vp, C.c_int, dp, ip]
        L.oracle_sim_nblocks.argtypes = [vp]
        L.oracle_sim_block_loc.argtypes = [vp, C.c_int, ip]
        L.oracle_sim_diagnostics.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.oracle_metric_at.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp, dp, dp]
        L.oracle_christoffel.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp]
        L.oracle_state.argtypes = [dp, dp, dp, dp]
        L.oracle_flux.argtypes = [C.c_double, dp, dp, dp, C.c_double, C.c_int, dp]
        L.oracle_fms.argtypes = [C.c_double, dp, dp, dp, C.c_int, dp]
        L.oracle_recon.argtypes = [C.c_int, dp, dp]
        L.oracle_c2p.argtypes = [C.c_int, dp, dp, C.c_double, dp, dp, ip]
        L.oracle_kastaun_residual.argtypes = [dp, C.c_double, C.c_double, dp, dp, dp]
        L.oracle_vsq_residual.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
        L.oracle_plant.argtypes = [C.c_int]
        L.oracle_heating.restype = C.c_double
        L.oracle_heating.argtypes = [C.POINTER(OracleCfg), dp]
        L.oracle_heating_fraction.restype = C.c_double
        L.oracle_heating_fraction.argtypes = [C.POINTER(OracleCfg), dp]
        L.oracle_physical_metric.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp]
        _lib = L
    return _lib


def make_cfg(d):
    """Fill an OracleCfg from a run-configuration dict."""
    c = OracleCfg()
    for name, ctype in OracleCfg._fields_:
        if name not in d:
            continue
        v = d[name]
        if name == "bc":
            for i in range(3):
                for s in range(2):
                    c.bc[i][s] = int(v[i][s])
        elif name in ("nx", "nb", "xmin", "xmax"):
            arr = getattr(c, name)
            for i in range(3):
                arr[i] = v[i]
        elif name in ("nref", "ref_level", "ref_min", "ref_max"):
            raise ValueError("give refinement regions as cfg['refine']")
        else:
            setattr(c, name, v)
    regions = d.get("refine", ())
    if len(regions) > MAX_REF:
        raise ValueError(f"at most {MAX_REF} refinement regions")
    c.nref = len(regions)
    for r, (level, lo, hi) in enumerate(regions):
        c.ref_level[r] = int(level)
        for i in range(3):
            c.ref_min[r][i] = float(lo[i])
            c.ref_max[r][i] = float(hi[i])
    return c


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


FIELDS = dict(prim=0, cons=1, flux1=2, flux2=3, flux3=4, alfven=5, emf=6, flag=7, mask=8,
              c2p_status=9, c2p_iters=10, flag_recovery=11, dudt=12)
TASKS = dict(fix_primitive=0, conservative=1, fluxes=2, ct=3, flux_div=4, update=5, source=6,
             recovery=7, heating=8, fix_recovery=9, exchange=10, bc=11, timestep=12,
             flux_correction=13)


DIAG_NAMES = ("mdot", "edot", "jdot", "phi", "mass", "energy", "divb_max", "ub_max")

# Operation counting (liboracle_count.so, counting.h): task rows follow the
# ORT_* order, operation columns the OpKind order.
TASK_NAMES = ("fix_primitive", "conservative", "fluxes", "ct", "flux_div", "update", "source",
              "recovery", "heating", "fix_recovery", "exchange", "bc", "timestep", "flux_correction")
OP_NAMES = ("add", "mul", "div", "sqrt", "cmp", "minmax", "fabs", "exp", "log", "pow", "sincos")
_clib = None


def lib_count():
    """The operation-counting build of the oracle (single-threaded)."""
    global _clib
    if _clib is None:
        path = os.path.join(_HERE, "_build", "liboracle_count.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        L.oracle_sim_create.restype = vp
        L.oracle_sim_create.argtypes = [C.POINTER(OracleCfg)]
        L.oracle_sim_destroy.argtypes = [vp]
        L.oracle_sim_init.argtypes = [vp]
        L.oracle_sim_step.argtypes = [vp, C.c_int]
        L.oracle_op_counts.argtypes = [C.POINTER(C.c_uint64)]
        L.oracle_op_reset.argtypes = []
        L.oracle_sim_field_size.restype = C.c_int64
        L.oracle_sim_field_size.argtypes = [vp, C.c_int]
        L.oracle_sim_get.argtypes = [vp, C.c_int, C.c_int, vp]
        _clib = L
    return _clib


def op_counts(cfg_dict, cycles=1, return_prim=False):
    """Operations performed by each task over `cycles` RK2 cycles after the
    initial data, as a (task, operation) array (TASK_NAMES x OP_NAMES); with
    return_prim, also the flat primitive arena of the base container."""
    L = lib_count()
    d = dict(cfg_dict)
    d["nthreads"] = 1
    cfg = make_cfg(d)
    p = L.oracle_sim_create(C.byref(cfg))
    if not p:
        raise ValueError("oracle_sim_create rejected the configuration")
    try:
        assert L.oracle_sim_init(p) == 0
        L.oracle_op_reset()
        assert L.oracle_sim_step(p, cycles) == 0
        out = np.zeros(len(TASK_NAMES) * len(OP_NAMES), dtype=np.uint64)
        assert L.oracle_op_counts(out.ctypes.data_as(C.POINTER(C.c_uint64))) == 0
        prim = None
        if return_prim:
            prim = np.empty(L.oracle_sim_field_size(p, FIELDS["prim"]) // 8)
            assert L.oracle_sim_get(p, FIELDS["prim"], 0, prim.ctypes.data_as(C.c_void_p)) == 0
    finally:
        L.oracle_sim_destroy(p)
    counts = out.reshape(len(TASK_NAMES), len(OP_NAMES))
    return (counts, prim) if return_prim else counts


class OracleSim:
    """One oracle run: Parthenon-like blocks, stage containers, RK2 cycles."""

    def __init__(self, cfg_dict, nthreads=0):
        d = dict(cfg_dict)
        d.setdefault("nthreads", nthreads)
        self._cfg = make_cfg(d)
        self._p = lib().oracle_sim_create(C.byref(self._cfg))
        if not self._p:
            raise ValueError("oracle_sim_create rejected the configuration")
        self.cfg = cfg_dict
        self.nx = list(cfg_dict["nx"])
        self.nb = list(cfg_dict["nb"])
        self.nblocks = lib().oracle_sim_nblocks(self._p)
        xd = np.zeros(6)
        n = np.zeros(9, dtype=np.int32)
        lib().oracle_sim_block_geom(self._p, 0, _dp(xd), n.ctypes.data_as(C.POINTER(C.c_int)))
        self.n = [int(n[0]), int(n[1]), int(n[2])]
        self.s = [int(n[3]), int(n[4]), int(n[5])]
        self.e = [int(n[6]), int(n[7]), int(n[8])]

    def __del__(self):
        if getattr(self, "_p", None):
            lib().oracle_sim_destroy(self._p)
            self._p = None

    def block_loc(self, b):
        """(level, lx1, lx2, lx3) of block b."""
        out = np.zeros(4, dtype=np.int32)
        assert lib().oracle_sim_block_loc(self._p, b, out.ctypes.data_as(C.POINTER(C.c_int))) == 0
        return tuple(int(x) for x in out)

    def block_geom(self, b):
        xd = np.zeros(6)
        n = np.zeros(9, dtype=np.int32)
        lib().oracle_sim_block_geom(self._p, b, _dp(xd), n.ctypes.data_as(C.POINTER(C.c_int)))
        return xd, n

    def init(self):
        assert lib().oracle_sim_init(self._p) == 0

    def diagnostics(self, shell, container=0, magnitudes=False):
        """Shell integrals at interior x1 index `shell`, volume integrals and
        div B of one container (oracle/diagnostics.cpp), as a dict; with
        magnitudes, also the dict of sums of |contribution| per entry."""
        out = np.zeros(len(DIAG_NAMES))
        mag = np.zeros(len(DIAG_NAMES))
        assert lib().oracle_sim_diagnostics(self._p, container, shell, _dp(out), _dp(mag)) == 0
        res = dict(zip(DIAG_NAMES, out))
        return (res, dict(zip(DIAG_NAMES, mag))) if magnitudes else res

    def step(self, n=1):
        lib().oracle_sim_step(self._p, n)

    def task(self, name, c0=0, c1=0, beta_dt=0.0):
        out = C.c_double(0.0)
        r = lib().oracle_sim_task(self._p, TASKS[name], c0, c1, beta_dt, C.byref(out))
        assert r == 0
        return out.value

    @property
    def time(self):
        return lib().oracle_sim_time(self._p)

    @property
    def dt(self):
        return lib().oracle_sim_dt(self._p)

    def _dtype(self, field):
        if field == "mask":
            return np.uint8
        if field in ("c2p_status", "c2p_iters"):
            return np.int32
        return np.float64

    def get(self, field, container=0):
        """Whole ghosted array shaped (nblocks, ncomp, nk, nj, ni)."""
        fid = FIELDS[field]
        nbytes = lib().oracle_sim_field_size(self._p, fid)
        dt = np.dtype(self._dtype(field))
        a = np.empty(nbytes // dt.itemsize, dtype=dt)
        lib().oracle_sim_get(self._p, fid, container, a.ctypes.data_as(C.c_void_p))
        ncell = self.n[0] * self.n[1] * self.n[2]
        return a.reshape(self.nblocks, -1, self.n[2], self.n[1], self.n[0]) if a.size % (
            self.nblocks * ncell) == 0 else a

    def set(self, field, arr, container=0):
        fid = FIELDS[field]
        a = np.ascontiguousarray(arr, dtype=self._dtype(field))
        assert a.nbytes == lib().oracle_sim_field_size(self._p, fid)
        lib().oracle_sim_set(self._p, fid, container, a.ctypes.data_as(C.c_void_p))

    def interior(self, arr):
        s, e = self.s, self.e
        return arr[..., s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1]


PLANTS = dict(none=0, conn_index=1, shift_sign=2, ppm_term=3, kastaun_qbar=4, ct_sign=5,
              emt_pressure=6, no_flux_correction=7, prolong_slope=8,
              cks_jacobian=9, cks_curl=10, heat_howes=11, heat_kawazura=12,
              heat_werner=13, heat_rowan=14, heat_sharma=15, bl_gthth=16)


def plant(name):
    """Activate a planted mistake in the oracle (tests of the pins only)."""
    lib().oracle_plant(PLANTS[name])


# ---- pointwise functions ------------------------------------------------
def metric_at(kind, x, a=0.0, h=0.0):
    x = np.asarray(x, dtype=np.float64)
    gcov = np.zeros(16)
    gcon = np.zeros(16)
    gdet = C.c_double(0.0)
    ok = lib().oracle_metric_at(kind, _dp(x), a, h, _dp(gcov), _dp(gcon), C.byref(gdet))
    return gcov.reshape(4, 4), gcon.reshape(4, 4), gdet.value, ok


def physical_metric(kind, x, a=0.0, h=0.0):
    x = np.asarray(x, dtype=np.float64)
    g = np.zeros(16)
    lib().oracle_physical_metric(kind, _dp(x), a, h, _dp(g))
    return g.reshape(4, 4)


def christoffel(kind, x, a=0.0, h=0.0):
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros(64)
    lib().oracle_christoffel(kind, _dp(x), a, h, _dp(out))
    return out.reshape(4, 4, 4)


def state(prim, gcov, gcon):
    out = np.zeros(17)
    p = np.ascontiguousarray(prim, dtype=np.float64)
    gc = np.ascontiguousarray(gcov, dtype=np.float64)
    gn = np.ascontiguousarray(gcon, dtype=np.float64)
    lib().oracle_state(_dp(p), _dp(gc), _dp(gn), _dp(out))
    return dict(ucon=out[0:4], ucov=out[4:8], bcon=out[8:12], bcov=out[12:16], bsq=out[16])


def flux(gamma, prim, gcov, gcon, gdet, direction):
    out = np.zeros(10)
    p = np.ascontiguousarray(prim, dtype=np.float64)
    gc = np.ascontiguousarray(gcov, dtype=np.float64)
    gn = np.ascontiguousarray(gcon, dtype=np.float64)
    lib().oracle_flux(gamma, _dp(p), _dp(gc), _dp(gn), gdet, direction, _dp(out))
    return out


def fms(gamma, prim, gcov, gcon, direction):
    out = np.zeros(2)
    p = np.ascontiguousarray(prim, dtype=np.float64)
    gc = np.ascontiguousarray(gcov, dtype=np.float64)
    gn = np.ascontiguousarray(gcon, dtype=np.float64)
    lib().oracle_fms(gamma, _dp(p), _dp(gc), _dp(gn), direction, _dp(out))
    return out


def recon(limiter, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros(2)
    lib().oracle_recon(limiter, _dp(q), _dp(out))
    return out


def c2p(scheme, U, prim_guess, gamma, gcov, gcon):
    """scheme: 0 2D, 1 1Dvsq, 2 1D, 3 Kastaun.  Returns (status, prim, iters)."""
    Uc = np.ascontiguousarray(U, dtype=np.float64)
    p = np.array(prim_guess, dtype=np.float64)
    gc = np.ascontiguousarray(gcov, dtype=np.float64)
    gn = np.ascontiguousarray(gcon, dtype=np.float64)
    it = C.c_int(0)
    st = lib().oracle_c2p(scheme, _dp(Uc), _dp(p), gamma, _dp(gc), _dp(gn), C.byref(it))
    return st, p, it.value


def kastaun_residual(Ucont, mu, gamma, gcov, gcon):
    """|f(mu)| of Kastaun's master function for the container-form conserved
    vector Ucont (8 values); None where the state has no invariants."""
    U = np.ascontiguousarray(Ucont[:8], dtype=np.float64)
    gc = np.ascontiguousarray(gcov, dtype=np.float64)
    gn = np.ascontiguousarray(gcon, dtype=np.float64)
    out = C.c_double(0.0)
    r = lib().oracle_kastaun_residual(_dp(U), float(mu), gamma, _dp(gc), _dp(gn), C.byref(out))
    return out.value if r == 0 else None


def vsq_residual(Ucont, K, vsq, gamma, gcov, gcon):
    """(f, scale) of the 1Dvsq residual at v^2 = vsq with entropy K for the
    container-form conserved vector Ucont (8 values); None when D <= 0 or
    vsq is outside [0, 1)."""
    U = np.ascontiguousarray(Ucont[:8], dtype=np.float64)
    gc = np.ascontiguousarray(gcov, dtype=np.float64)
    gn = np.ascontiguousarray(gcon, dtype=np.float64)
    res, sc = C.c_double(0.0), C.c_double(0.0)
    r = lib().oracle_vsq_residual(_dp(U), float(K), float(vsq), gamma, _dp(gc), _dp(gn), C.byref(res),
                                  C.byref(sc))
    return (res.value, sc.value) if r == 0 else None


def heating(cfg_dict, cell):
    c = make_cfg(cfg_dict)
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    return lib().oracle_heating(C.byref(c), _dp(cell))


def heating_fraction(cfg_dict, cell):
    c = make_cfg(cfg_dict)
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    return lib().oracle_heating_fraction(C.byref(c), _dp(cell))
