"""ctypes binding of the pangu_sim host driver (include/pangu_sim.h).

Device memory, kernels and streams all live in libpangu_hip.so; this module
moves configuration in and fields out.  It raises if the library is missing
or cannot be loaded: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import lib_path


MAX_REF = 8  # PANGU_SIM_MAX_REF


def refine_fields(c, regions):
    """Static refinement regions [(level, (x1min, x2min, x3min), (x1max, x2max,
    x3max)), ...] into the nref / ref_level / ref_min / ref_max fields of a
    configuration struct."""
    if len(regions) > MAX_REF:
        raise ValueError(f"at most {MAX_REF} refinement regions")
    c.nref = len(regions)
    for r, (level, lo, hi) in enumerate(regions):
        c.ref_level[r] = int(level)
        for d in range(3):
            c.ref_min[r][d] = float(lo[d])
            c.ref_max[r][d] = float(hi[d])


class SimCfg(C.Structure):
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


FIELDS = dict(prim=0, cons=1, flux1=2, flux2=3, flux3=4, alfven=5, emf=6, flag=7, mask=8,
              c2p_status=9, c2p_iters=10, flag_recovery=11, dudt=12, c2p_root=13)
TASKS = dict(fix_primitive=0, conservative=1, fluxes=2, ct=3, flux_div=4, update=5, source=6,
             recovery=7, heating=8, fix_recovery=9, exchange=10, bc=11, timestep=12,
             flux_correction=13)
MODES = dict(tasks=0, fused=1, graph=2)
# pangu_diagnostics outputs, in PANGU_DIAG_* order
DIAG_NAMES = ("mdot", "edot", "jdot", "phi", "mass", "energy", "divb_max", "ub_max")
KERNEL_FAMILIES = dict(fix_cons=0, flux_x1=1, flux_x2=2, flux_x3=3, ct=4, div_update=5, recovery=6)

_lib = None


def lib():
    """Load libpangu_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python -m pangu_amd.build` first")
        L = C.CDLL(path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.pangu_sim_create.argtypes = [C.POINTER(SimCfg), C.c_int, C.POINTER(vp)]
        L.pangu_sim_create_batch.argtypes = [C.POINTER(SimCfg), C.c_int, C.c_int, C.POINTER(vp)]
        L.pangu_sim_mesh_state.argtypes = [vp, C.c_int, dp, dp, ip, ip]
        L.pangu_sim_diagnostics.argtypes = [vp, C.c_int, ip, dp]
        L.pangu_sim_context.argtypes = [vp, C.POINTER(vp), ip, C.POINTER(vp)]
        L.pangu_sim_record_roots.argtypes = [vp, C.c_int]
        L.pangu_sim_set_pair_flux.argtypes = [vp, C.c_int]
        L.pangu_sim_set_flux_march.argtypes = [vp, C.c_int, C.c_int]
        L.pangu_sim_set_c2p_gate.argtypes = [vp, C.c_int]
        L.pangu_sim_nblocks.argtypes = [vp]
        L.pangu_sim_plain_loop.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, dp]
        L.pangu_sim_autotune.argtypes = [vp, C.c_int, C.c_int, ip, ip, ip, ip, ip, ip, ip]
        L.pangu_sim_set_persistent_recovery.argtypes = [vp, C.c_int]
        L.pangu_sim_set_tile_size.argtypes = [vp, C.c_int, C.c_int]
        L.pangu_sim_destroy.argtypes = [vp]
        L.pangu_sim_init.argtypes = [vp]
        L.pangu_sim_step.argtypes = [vp, C.c_int, C.c_int]
        L.pangu_sim_sync.argtypes = [vp]
        L.pangu_sim_timed_cycles.argtypes = [vp, C.c_int, C.c_int, dp]
        L.pangu_sim_time_phases.argtypes = [vp, C.c_int, dp, C.c_int]
        L.pangu_sim_phase_name.restype = C.c_char_p
        L.pangu_sim_phase_name.argtypes = [C.c_int]
        L.pangu_sim_task.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, dp]
        L.pangu_sim_field_size.restype = C.c_int64
        L.pangu_sim_field_size.argtypes = [vp, C.c_int]
        L.pangu_sim_get.argtypes = [vp, C.c_int, C.c_int, vp]
        L.pangu_sim_set.argtypes = [vp, C.c_int, C.c_int, vp]
        L.pangu_sim_time.restype = C.c_double
        L.pangu_sim_time.argtypes = [vp]
        L.pangu_sim_dt.restype = C.c_double
        L.pangu_sim_dt.argtypes = [vp]
        L.pangu_sim_block_geom.argtypes = [vp, C.c_int, dp, ip]
        L.pangu_sim_block_loc.argtypes = [vp, C.c_int, ip]
        L.pangu_sim_set_tile_order.argtypes = [vp, C.c_int, C.c_int]
        L.pangu_sim_set_split_source.argtypes = [vp, C.c_int]
        L.pangu_sim_set_fuse_ct.argtypes = [vp, C.c_int]
        L.pangu_sim_set_overlap_exchange.argtypes = [vp, C.c_int]
        L.pangu_sim_set_fuse_update_recovery.argtypes = [vp, C.c_int]
        L.pangu_sim_set_fuse_fix_x1.argtypes = [vp, C.c_int]
        L.pangu_sim_set_share_reconstruction.argtypes = [vp, C.c_int]
        L.pangu_sim_time_kernels.argtypes = [vp, C.c_int, dp, ip, C.c_int]
        L.pangu_sim_kernel_name.restype = C.c_char_p
        L.pangu_sim_kernel_name.argtypes = [vp, C.c_int]
        L.pangu_sim_copy_calibration.argtypes = [vp, C.c_int64, C.c_int, dp]
        L.pangu_sim_last_error.restype = C.c_char_p
        L.pangu_sim_last_error.argtypes = [vp]
        _lib = L
    return _lib


def make_cfg(d):
    c = SimCfg()
    for name, _ in SimCfg._fields_:
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
    refine_fields(c, d.get("refine", ()))
    return c


class HipSim:
    """One run of the HIP path on a GPU.  Given a list of configurations it
    runs them as one batched ensemble (pangu_sim_create_batch)."""

    def __init__(self, cfg_dict, device=0):
        cfgs = list(cfg_dict) if isinstance(cfg_dict, (list, tuple)) else [cfg_dict]
        self.cfg = cfgs[0]
        self.cfgs = cfgs
        self.nmesh = len(cfgs)
        arr = (SimCfg * self.nmesh)()
        for m, d in enumerate(cfgs):
            arr[m] = make_cfg(d)
        self._cfg = arr
        p = C.c_void_p()
        r = lib().pangu_sim_create_batch(arr, self.nmesh, device, C.byref(p))
        if r != 0 or not p:
            raise RuntimeError(f"pangu_sim_create failed ({r})")
        self._p = p
        self.nb = list(cfgs[0]["nb"])
        self.nblocks = lib().pangu_sim_nblocks(self._p)
        self.shapes, self.starts, self.ends = [], [], []
        for b in range(self.nblocks):
            _, n = self.block_geom(b)
            self.shapes.append((int(n[0]), int(n[1]), int(n[2])))
            self.starts.append([int(n[3]), int(n[4]), int(n[5])])
            self.ends.append([int(n[6]), int(n[7]), int(n[8])])
        self.uniform = len(set(self.shapes)) == 1
        self.n = list(self.shapes[0])
        self.s = self.starts[0]
        self.e = self.ends[0]

    def mesh_state(self, m):
        """(dt, time, first block, block count) of mesh m."""
        dt, tm = C.c_double(), C.c_double()
        fb, nb = C.c_int(), C.c_int()
        self._check(lib().pangu_sim_mesh_state(self._p, m, C.byref(dt), C.byref(tm), C.byref(fb),
                                               C.byref(nb)), "mesh_state")
        return dt.value, tm.value, fb.value, nb.value

    def diagnostics(self, shell, container=0):
        """Shell fluxes at interior x1 index `shell` (one per mesh, or one for
        all), volume integrals and div B of a container (pangu_diagnostics):
        a dict of DIAG_NAMES for a one-mesh simulation, else a list of them."""
        nm = self.nmesh
        sh = np.asarray([shell] * nm if np.isscalar(shell) else shell, dtype=np.int32)
        out = np.zeros(nm * len(DIAG_NAMES))
        self._check(lib().pangu_sim_diagnostics(self._p, container, sh.ctypes.data_as(C.POINTER(C.c_int)),
                                                out.ctypes.data_as(C.POINTER(C.c_double))), "diagnostics")
        res = [dict(zip(DIAG_NAMES, out[m * len(DIAG_NAMES):(m + 1) * len(DIAG_NAMES)])) for m in range(nm)]
        return res[0] if nm == 1 else res

    def record_roots(self, on=True):
        """Kastaun recovery records each cell's root (get("c2p_root"))."""
        self._check(lib().pangu_sim_record_roots(self._p, int(on)), "record_roots")

    def context(self):
        """(pangu_ctx pointer, container handles (base, stage 1), stream) of
        this simulation, for calling pangu_hip entry points directly."""
        ctx, st = C.c_void_p(), C.c_void_p()
        h = (C.c_int * 2)()
        self._check(lib().pangu_sim_context(self._p, C.byref(ctx), h, C.byref(st)), "context")
        return ctx, (h[0], h[1]), st

    def history(self, cycles, every, shell, mode="graph"):
        """Steps `cycles` cycles and records the diagnostics every `every`
        cycles: a list of (cycles done, time of mesh 0, diagnostics)."""
        rows, done = [], 0
        while done < cycles:
            n = min(every, cycles - done)
            self.step(n, mode)
            done += n
            rows.append((done, self.mesh_state(0)[1], self.diagnostics(shell)))
        return rows

    def autotune(self, cycles=3, use_cache=True):
        """Runs the autotuner (state restored afterwards); returns the chosen
        tile order per kernel family and the source split."""
        orders = (C.c_int * len(KERNEL_FAMILIES))()
        split, share, pair = C.c_int(), C.c_int(), C.c_int()
        march = (C.c_int * 2)()
        tiles = (C.c_int * len(KERNEL_FAMILIES))()
        persist = C.c_int()
        self._check(lib().pangu_sim_autotune(self._p, cycles, int(use_cache), orders, C.byref(split),
                                             C.byref(share), C.byref(pair), march, tiles,
                                             C.byref(persist)), "autotune")
        names = sorted(KERNEL_FAMILIES, key=KERNEL_FAMILIES.get)
        return {"order": {n: int(orders[KERNEL_FAMILIES[n]]) for n in names},
                "split_source": split.value, "share_reconstruction": share.value,
                "pair_flux": pair.value, "flux_march": [march[0], march[1]],
                "tile": {n: int(tiles[KERNEL_FAMILIES[n]]) for n in names},
                "persistent_recovery": persist.value}

    def close(self):
        if getattr(self, "_p", None):
            lib().pangu_sim_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def _check(self, r, what):
        if r != 0:
            msg = lib().pangu_sim_last_error(self._p)
            raise RuntimeError(f"{what} failed ({r}): {msg.decode() if msg else ''}")

    def block_geom(self, b):
        xd = np.zeros(6)
        n = np.zeros(9, dtype=np.int32)
        lib().pangu_sim_block_geom(self._p, b, xd.ctypes.data_as(C.POINTER(C.c_double)),
                                   n.ctypes.data_as(C.POINTER(C.c_int)))
        return xd, n

    def block_loc(self, b):
        """(level, lx1, lx2, lx3) of block b."""
        out = np.zeros(4, dtype=np.int32)
        self._check(lib().pangu_sim_block_loc(self._p, b, out.ctypes.data_as(C.POINTER(C.c_int))),
                    "block_loc")
        return tuple(int(x) for x in out)

    def init(self):
        self._check(lib().pangu_sim_init(self._p), "init")

    def step(self, n=1, mode="tasks"):
        self._check(lib().pangu_sim_step(self._p, n, MODES[mode]), "step")
        self._check(lib().pangu_sim_sync(self._p), "sync")

    def timed_cycles(self, n, mode="graph"):
        ms = C.c_double(0.0)
        self._check(lib().pangu_sim_timed_cycles(self._p, n, MODES[mode], C.byref(ms)), "timed")
        return ms.value

    def time_phases(self, reps=5):
        out = np.zeros(16)
        n = lib().pangu_sim_time_phases(self._p, reps, out.ctypes.data_as(C.POINTER(C.c_double)), 16)
        if n < 0:
            self._check(n, "time_phases")
        return {lib().pangu_sim_phase_name(i).decode(): float(out[i]) for i in range(n)}

    def time_kernels(self, ncycles=3):
        """{kernel: (launches per cycle, ms per cycle)} from HIP events
        recorded after every launch of ncycles fused cycles."""
        ms = np.zeros(64)
        calls = np.zeros(64, dtype=np.int32)
        n = lib().pangu_sim_time_kernels(self._p, ncycles, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                         calls.ctypes.data_as(C.POINTER(C.c_int)), 64)
        if n < 0:
            self._check(n, "time_kernels")
        return {lib().pangu_sim_kernel_name(self._p, i).decode(): (int(calls[i]), float(ms[i]))
                for i in range(min(n, 64))}

    def set_tile_order(self, family, order):
        """Tile order of one kernel family (KERNEL_FAMILIES; 0 plane-major,
        1 column-major).  Results are unchanged; only speed differs."""
        fam = KERNEL_FAMILIES[family] if isinstance(family, str) else int(family)
        self._check(lib().pangu_sim_set_tile_order(self._p, fam, int(order)), "set_tile_order")

    def set_split_source(self, split):
        """Fused stage: geometric source in a kernel of its own (1) or fused
        with the flux divergence and update (0).  Results are unchanged."""
        self._check(lib().pangu_sim_set_split_source(self._p, int(split)), "set_split_source")

    def set_pair_flux(self, dir_mask):
        """Bit d-1: the sweep along d runs two lanes per face (k_flux_pair).
        Results are unchanged."""
        self._check(lib().pangu_sim_set_pair_flux(self._p, int(dir_mask)), "set_pair_flux")

    def set_persistent_recovery(self, on):
        """Recovery as a persistent kernel pulling tiles from a queue (1) or
        one workgroup per tile (0).  Results are unchanged."""
        self._check(lib().pangu_sim_set_persistent_recovery(self._p, int(on)), "set_persistent_recovery")

    def set_tile_size(self, family, cells):
        """Cells per tile (workgroup size) of a kernel family's launches:
        256, 128 or 64.  Results are unchanged."""
        self._check(lib().pangu_sim_set_tile_size(self._p, KERNEL_FAMILIES[family], int(cells)),
                    "set_tile_size")

    def set_flux_march(self, direction, depth):
        """The sweep along `direction` (2 or 3) runs `depth` faces per thread,
        one reconstruction per cell (0: one face per thread).  Results are
        unchanged."""
        self._check(lib().pangu_sim_set_flux_march(self._p, int(direction), int(depth)), "set_flux_march")

    def set_fuse_ct(self, fuse):
        """Fused stage: constrained-transport face B fluxes formed inside the
        update kernel (1, default) or written by their own kernel first (0).
        Results are unchanged."""
        self._check(lib().pangu_sim_set_fuse_ct(self._p, int(fuse)), "set_fuse_ct")

    def set_fuse_fix_x1(self, on):
        """Fused stage: FixPrimitive + conservative and the X1 sweep in one
        kernel (1) or two (0, default).  Results are unchanged."""
        self._check(lib().pangu_sim_set_fuse_fix_x1(self._p, int(on)), "set_fuse_fix_x1")

    def set_fuse_update_recovery(self, on):
        """Fused stage: flux-divergence update and recovery in one kernel (1,
        default) or two (0).  Results are unchanged."""
        self._check(lib().pangu_sim_set_fuse_update_recovery(self._p, int(on)), "set_fuse_update_recovery")

    def set_overlap_exchange(self, on):
        """Fused stage: the ghost exchange after the update runs on a second
        stream from the start of the stage (1) or on the stage's stream after
        the update (0); -1 (default): 1 on a refined mesh only.  Results are
        unchanged."""
        self._check(lib().pangu_sim_set_overlap_exchange(self._p, int(on)), "set_overlap_exchange")

    def set_share_reconstruction(self, share):
        """X1 sweep: one reconstruction per cell shared between its two faces
        (1) or one per face side (0).  Results are unchanged."""
        self._check(lib().pangu_sim_set_share_reconstruction(self._p, int(share)), "share_reconstruction")

    def copy_calibration(self, n, reps=10):
        ms = C.c_double(0.0)
        self._check(lib().pangu_sim_copy_calibration(self._p, n, reps, C.byref(ms)), "copy")
        return ms.value

    def task(self, name, c0=0, c1=0, beta_dt=0.0):
        out = C.c_double(0.0)
        self._check(lib().pangu_sim_task(self._p, TASKS[name], c0, c1, beta_dt, C.byref(out)), name)
        return out.value

    @property
    def time(self):
        return lib().pangu_sim_time(self._p)

    @property
    def dt(self):
        return lib().pangu_sim_dt(self._p)

    def _dtype(self, field):
        if field == "mask":
            return np.uint8
        if field in ("c2p_status", "c2p_iters"):
            return np.int32
        return np.float64

    def get(self, field, container=0):
        """Whole arena of a field: (nblocks, ncomp, nk, nj, ni) when all
        blocks share a shape, else the flat arena (see blocks())."""
        fid = FIELDS[field]
        nbytes = lib().pangu_sim_field_size(self._p, fid)
        dt = np.dtype(self._dtype(field))
        a = np.empty(nbytes // dt.itemsize, dtype=dt)
        self._check(lib().pangu_sim_get(self._p, fid, container, a.ctypes.data_as(C.c_void_p)), "get")
        if self.uniform:
            return a.reshape(self.nblocks, -1, self.n[2], self.n[1], self.n[0])
        return a

    def blocks(self, flat):
        """Per-block (ncomp, nk, nj, ni) views of a flat arena."""
        cells = [s[0] * s[1] * s[2] for s in self.shapes]
        ncomp = flat.size // sum(cells)
        out, off = [], 0
        for s, c in zip(self.shapes, cells):
            out.append(flat[off:off + ncomp * c].reshape(ncomp, s[2], s[1], s[0]))
            off += ncomp * c
        return out

    def block_interior(self, arr, b):
        """Interior of block b's (ncomp, nk, nj, ni) array."""
        s, e = self.starts[b], self.ends[b]
        return arr[..., s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1]

    def set(self, field, arr, container=0):
        fid = FIELDS[field]
        a = np.ascontiguousarray(arr, dtype=self._dtype(field))
        assert a.nbytes == lib().pangu_sim_field_size(self._p, fid)
        self._check(lib().pangu_sim_set(self._p, fid, container, a.ctypes.data_as(C.c_void_p)), "set")

    def interior(self, arr):
        s, e = self.s, self.e
        return arr[..., s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1]


def plain_loop(sims, ncycles, mode="graph"):
    """Milliseconds for ncycles cycles of every simulation run in turn on one
    stream (the loop around the one-mesh API that batching replaces)."""
    arr = (C.c_void_p * len(sims))(*[s._p.value for s in sims])
    ms = C.c_double()
    r = lib().pangu_sim_plain_loop(arr, len(sims), ncycles, MODES[mode], C.byref(ms))
    if r != 0:
        raise RuntimeError(f"plain_loop failed ({r})")
    return ms.value
