"""A Parthenon-like host driver over the task-level C ABI, in Python.

It drives the entry points of include/pangu_hip.h (through the ctypes
mirror in pangu_amd.abi) the way PANGU's task list drives its tasks
(pangu/src/task_list/ideal_grmhd.cc:23-120): per RK stage FixPrimitive,
CalculateConservative, CalculateFluxes, ConstraintedTransport, flux
correction, FluxDivergence, UpdateIndependentData, the ghost exchange,
AddGeometricSource, Recovery, ApplyElectronHeating, FixRecovery, the
physical boundaries and, after the last stage, EstimateTimestep and the dt
control (ideal_grmhd.cc:29 for the RK2 weights).

Unlike the package's own driver (sim.hip, one arena per field), this one
lays its device memory out the way a Parthenon MeshBlockPack does: one
allocation per (meshblock, variable), with fluxes per container.  It is the
reference point for an adapter on the Parthenon side (see the integration
notes) and shows the boundary taking caller-owned buffers of any layout.
Initial data are copied in from the host.
"""
import ctypes as C

import numpy as np

from . import abi


class HipRuntime:
    """The HIP runtime that libpangu_hip.so links (loaded with it)."""

    def __init__(self):
        abi.lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [C.c_int]
        self.h = h

    def check(self, r, what):
        if r != 0:
            raise RuntimeError(f"{what} failed (hipError {r})")

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(nbytes, 1)), "hipMalloc")
        self.check(self.h.hipMemset(p, 0, max(nbytes, 1)), "hipMemset")
        return p.value

    def free(self, p):
        self.h.hipFree(C.c_void_p(p))

    def upload(self, p, arr):
        a = np.ascontiguousarray(arr)
        self.check(self.h.hipMemcpy(C.c_void_p(p), a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "hipMemcpy H2D")

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        self.check(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p), out.nbytes, 2),
                   "hipMemcpy D2H")
        return out

    def sync(self):
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")


def _table(ptrs, ctype=C.c_double):
    """A host array of device pointers, as the ABI's pointer tables."""
    T = C.POINTER(ctype) * len(ptrs)
    return T(*[C.cast(C.c_void_p(p), C.POINTER(ctype)) for p in ptrs])


class TaskList:
    """One uniform mesh (a run-configuration dict of pangu_amd.configs),
    driven task by task through the C ABI on per-(block, variable) device
    buffers.  Containers: "base" (0), the stage container "1" and dU/dt."""

    FIELDS = ("prim", "cons")

    def __init__(self, cfg, device=0):
        self.cfg = cfg
        self.L = abi.lib()
        self.hip = HipRuntime()
        self.hip.check(self.hip.h.hipSetDevice(device), "hipSetDevice")
        self.blocks = abi.blocks_from(cfg)
        self.nb = len(self.blocks)
        B = self.blocks[0]
        self.n = [B.nx[d] + 2 * B.ng[d] for d in range(3)]
        self.s = [B.ng[d] for d in range(3)]
        self.e = [B.ng[d] + B.nx[d] - 1 for d in range(3)]
        self.cells = self.n[0] * self.n[1] * self.n[2]
        self.np = 5 + (3 if cfg["enable_B"] else 0) + 1 + (1 if cfg["enable_heating"] else 0)
        self.ndim = sum(1 for d in range(3) if cfg["nx"][d] > 1)
        self._cfg = abi.config_from(cfg)
        ctx = C.c_void_p()
        r = self.L.pangu_create(C.byref(self._cfg), self.blocks, self.nb, device, C.byref(ctx))
        if r != 0:
            raise RuntimeError(f"pangu_create failed ({r})")
        self.ctx = ctx
        self._owned = []
        nb, nc, cells = self.nb, self.np, self.cells

        def arrays(ncomp, itemsize=8):
            ptrs = [self.hip.alloc(cells * itemsize) for _ in range(nb * ncomp)]
            self._owned += ptrs
            return ptrs

        self.ptr = {}
        self.handle = {}
        for name in ("base", "1"):
            self.ptr[("prim", name)] = arrays(nc)
            self.ptr[("cons", name)] = arrays(nc)
            self.ptr[("flux", name)] = [arrays(nc) if d < self.ndim else None for d in range(3)]
        self.ptr[("cons", "dudt")] = arrays(nc)
        self._tables = []  # keep host tables alive while registered
        for name in ("base", "1", "dudt"):
            cont = abi.PanguContainer()
            tp = _table(self.ptr[("prim", "1" if name == "dudt" else name)])
            tu = _table(self.ptr[("cons", name)])
            cont.prim = C.cast(tp, abi._PPD)
            cont.cons = C.cast(tu, abi._PPD)
            self._tables += [tp, tu]
            for d in range(3):
                fl = self.ptr[("flux", name)][d] if name != "dudt" else None
                if fl is None:
                    cont.flux[d] = abi._PPD()
                else:
                    tf = _table(fl)
                    self._tables.append(tf)
                    cont.flux[d] = C.cast(tf, abi._PPD)
            h = C.c_int()
            self._ok(self.L.pangu_register_container(self.ctx, C.byref(cont), C.byref(h)), "register")
            self.handle[name] = h.value
        sh = abi.PanguShared()
        self.shared = {"alfven": arrays(3), "emf": arrays(3), "flag": arrays(1),
                       "flag_recovery": arrays(1), "mask": arrays(1, 1), "c2p_status": arrays(1, 4),
                       "c2p_iters": arrays(1, 4)}
        ctypes_of = {"mask": C.c_uint8, "c2p_status": C.c_int32, "c2p_iters": C.c_int32}
        for k, ptrs in self.shared.items():
            t = _table(ptrs, ctypes_of.get(k, C.c_double))
            self._tables.append(t)
            setattr(sh, k, C.cast(t, type(getattr(sh, k))))
        self._ok(self.L.pangu_register_shared(self.ctx, C.byref(sh)), "register_shared")
        self.dt = cfg["dt_init"]
        self.time = 0.0
        self._stage1_ready = False

    def _ok(self, r, what):
        if r != 0:
            msg = self.L.pangu_last_error(self.ctx)
            raise RuntimeError(f"{what} failed ({r}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "ctx", None):
            self.hip.sync()
            self.L.pangu_destroy(self.ctx)
            self.ctx = None
            for p in self._owned:
                self.hip.free(p)
            self._owned = []

    def __del__(self):
        self.close()

    # ------------------------------------------------------------- data
    def set(self, field, arr, container="base"):
        """arr: (nblocks, ncomp, nk, nj, ni), the ghosted arrays per block."""
        ptrs = self.ptr[(field, container)]
        nc = self.np
        for b in range(self.nb):
            for q in range(nc):
                self.hip.upload(ptrs[b * nc + q], np.ascontiguousarray(arr[b, q], dtype=np.float64))

    def get(self, field, container="base"):
        self.hip.sync()
        ptrs = self.ptr[(field, container)]
        nc = self.np
        out = np.empty((self.nb, nc, self.n[2], self.n[1], self.n[0]))
        for b in range(self.nb):
            for q in range(nc):
                out[b, q] = self.hip.download(ptrs[b * nc + q], (self.n[2], self.n[1], self.n[0]), np.float64)
        return out

    def interior(self, arr):
        s, e = self.s, self.e
        return arr[..., s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1]

    # ------------------------------------------------------------ tasks
    def stage(self, stage):
        """One RK stage at the reference's task granularity
        (ideal_grmhd.cc:52-116): stage 1 base -> "1" with beta 1/2, stage 2
        "1" -> base with beta 1; both update from the base container."""
        L, x, st = self.L, self.ctx, None
        h = self.handle
        c0, c1 = (h["base"], h["1"]) if stage == 1 else (h["1"], h["base"])
        bdt = (0.5 if stage == 1 else 1.0) * self.dt
        cfg = self.cfg
        self._ok(L.pangu_fix_primitive(x, c0, st), "fix_primitive")
        self._ok(L.pangu_calculate_conservative(x, c0, st), "calculate_conservative")
        self._ok(L.pangu_calculate_fluxes(x, c0, st), "calculate_fluxes")
        if cfg["enable_B"]:
            self._ok(L.pangu_constrained_transport(x, c0, st), "constrained_transport")
        self._ok(L.pangu_flux_correction(x, c0, st), "flux_correction")
        self._ok(L.pangu_flux_divergence(x, c0, h["dudt"], st), "flux_divergence")
        self._ok(L.pangu_update_independent(x, h["base"], h["dudt"], bdt, c1, st), "update_independent")
        if cfg["exchange_after"] == 0:
            self._ok(L.pangu_exchange_ghosts(x, c1, st), "exchange_ghosts")
        self._ok(L.pangu_add_geometric_source(x, c1, bdt, st), "add_geometric_source")
        self._ok(L.pangu_recovery(x, c1, st), "recovery")
        if cfg["enable_heating"]:
            self._ok(L.pangu_apply_electron_heating(x, c1, st), "apply_electron_heating")
        if cfg["recovery"] != abi.RECOVERY["robuster"]:
            self._ok(L.pangu_fix_recovery(x, c1, st), "fix_recovery")
        if cfg["exchange_after"] == 1:
            self._ok(L.pangu_exchange_ghosts(x, c1, st), "exchange_ghosts")
        self._ok(L.pangu_apply_boundary_conditions(x, c1, st), "apply_boundary_conditions")
        if stage == 2:
            est = (C.c_double * 1)()
            self._ok(L.pangu_estimate_timestep(x, c1, est, st), "estimate_timestep")
            return est[0]
        return None

    def cycle(self, n=1):
        """n RK2 cycles; dt_{n+1} = min(dt_factor dt_n, estimate)."""
        if not self._stage1_ready:
            if self.cfg.get("stage_init", 0) == 1:
                self.set("prim", self.get("prim", "base"), "1")
            self._stage1_ready = True
        for _ in range(n):
            self.stage(1)
            est = self.stage(2)
            self.time += self.dt
            self.dt = min(self.cfg["dt_factor"] * self.dt, est)
