"""ctypes binding of the learned Kastaun bracket (include/pangu_mlp.h).

Training, inference and evaluation run in libpangu_hip.so on the GPU; this
module only moves arrays.  It raises when the library is missing.
"""
import ctypes as C

import numpy as np

from . import sim as _sim

NPARAM = 5 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 + 5 + 5

_bound = False


def lib():
    global _bound
    L = _sim.lib()
    if not _bound:
        vp, fp, dp = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.pangu_mlp_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
        L.pangu_mlp_destroy.argtypes = [vp]
        L.pangu_mlp_last_error.restype = C.c_char_p
        L.pangu_mlp_last_error.argtypes = [vp]
        L.pangu_mlp_get_params.argtypes = [vp, fp]
        L.pangu_mlp_set_params.argtypes = [vp, fp]
        L.pangu_mlp_make_dataset.argtypes = [vp, C.c_int64, C.c_uint64, C.c_double, C.c_double,
                                             C.c_double, C.POINTER(C.c_int64)]
        L.pangu_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, fp]
        L.pangu_mlp_gradient.argtypes = [vp, fp, fp, C.c_int, fp, fp]
        L.pangu_mlp_predict.argtypes = [vp, fp, C.c_int, fp]
        L.pangu_mlp_evaluate.argtypes = [vp, C.c_double, dp]
        L.pangu_mlp_gemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp]
        L.pangu_sim_set_c2p_guess.argtypes = [vp, vp, C.c_double]
        L.pangu_sim_set_c2p_gate.argtypes = [vp, C.c_int]
        L.pangu_last_error.restype = C.c_char_p
        L.pangu_last_error.argtypes = [vp]
        L.pangu_mlp_harvest.argtypes = [vp, vp, C.c_int, C.c_int, C.POINTER(C.c_int64)]
        L.pangu_mlp_predict_wave.argtypes = [vp, fp, C.c_int, fp]
        L.pangu_mlp_set_kind.argtypes = [vp, C.c_int]
        _bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def gemm(A, B, device=0):
    """BF16 MFMA product of two host matrices with FP32 accumulation."""
    A, B = _f32(A), _f32(B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.zeros((M, N), dtype=np.float32)
    r = lib().pangu_mlp_gemm(device, M, N, K, _fp(A), _fp(B), _fp(Cm))
    if r != 0:
        raise RuntimeError(f"pangu_mlp_gemm failed ({r})")
    return Cm


class Mlp:
    """The 5-64-64-64-1 SiLU network that predicts Kastaun's root (kind 0)
    or the start of the 1Dvsq Newton iteration, log u^2 (kind 1)."""

    def __init__(self, seed=0, device=0, kind=0):
        p = C.c_void_p()
        r = lib().pangu_mlp_create(device, seed, C.byref(p))
        if r != 0 or not p:
            raise RuntimeError(f"pangu_mlp_create failed ({r})")
        self._p = p
        self.kind = kind
        self._check(lib().pangu_mlp_set_kind(self._p, int(kind)), "set_kind")

    def close(self):
        if getattr(self, "_p", None):
            lib().pangu_mlp_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def _check(self, r, what):
        if r != 0:
            msg = lib().pangu_mlp_last_error(self._p)
            raise RuntimeError(f"{what} failed ({r}): {msg.decode() if msg else ''}")

    def params(self):
        out = np.zeros(NPARAM, dtype=np.float32)
        self._check(lib().pangu_mlp_get_params(self._p, _fp(out)), "get_params")
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == NPARAM
        self._check(lib().pangu_mlp_set_params(self._p, _fp(p)), "set_params")

    def make_dataset(self, n, seed=1, a=0.9375, h=0.7, gamma=5.0 / 3.0):
        kept = C.c_int64()
        self._check(lib().pangu_mlp_make_dataset(self._p, n, seed, a, h, gamma, C.byref(kept)),
                    "make_dataset")
        return kept.value

    def train(self, epochs, batch=1024, lr=1e-3):
        loss = np.zeros(max(epochs, 1), dtype=np.float32)
        self._check(lib().pangu_mlp_train(self._p, epochs, batch, lr, _fp(loss)), "train")
        return loss[:epochs]

    def gradient(self, feats, target):
        feats, target = _f32(feats), _f32(target)
        n = target.size
        grad = np.zeros(NPARAM, dtype=np.float32)
        loss = C.c_float()
        self._check(lib().pangu_mlp_gradient(self._p, _fp(feats), _fp(target), n, C.byref(loss),
                                             _fp(grad)), "gradient")
        return loss.value, grad

    def predict(self, feats):
        feats = _f32(feats)
        n = feats.shape[0]
        out = np.zeros(n, dtype=np.float32)
        self._check(lib().pangu_mlp_predict(self._p, _fp(feats), n, _fp(out)), "predict")
        return out

    def predict_wave(self, feats):
        """The recovery kernel's in-wave inference path on raw features."""
        feats = _f32(feats)
        n = feats.shape[0]
        out = np.zeros(n, dtype=np.float32)
        self._check(lib().pangu_mlp_predict_wave(self._p, _fp(feats), n, _fp(out)), "predict_wave")
        return out

    def harvest(self, hipsim, container=0, stride=1):
        """Appends the states of a HipSim container (every stride-th interior
        cell, solved by plain Kastaun) to the data set; returns how many
        solved states were added."""
        ctx, handles, _ = hipsim.context()
        kept = C.c_int64()
        r = lib().pangu_mlp_harvest(self._p, ctx, handles[container], stride, C.byref(kept))
        if r != 0:
            raise RuntimeError(f"harvest failed ({r}): {lib().pangu_last_error(ctx).decode()}")
        return kept.value

    def evaluate(self, delta=0.02):
        out = np.zeros(7)
        self._check(lib().pangu_mlp_evaluate(self._p, delta, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "evaluate")
        keys = ("plain_iters", "learned_iters", "plain_evals", "learned_evals", "max_residual",
                "max_rel_diff", "failures")
        return dict(zip(keys, out.tolist()))

    def attach(self, hipsim, delta=0.02, gate=0):
        """Kastaun recovery of a HipSim starts from this model's bracket in
        the wavefronts holding a cell whose previous recovery took at least
        `gate` false-position steps (0: every cell)."""
        r = lib().pangu_sim_set_c2p_guess(hipsim._p, self._p, delta)
        if r != 0:
            raise RuntimeError(f"attach failed ({r})")
        r = lib().pangu_sim_set_c2p_gate(hipsim._p, int(gate))
        if r != 0:
            raise RuntimeError(f"gate failed ({r})")


def detach(hipsim):
    r = lib().pangu_sim_set_c2p_guess(hipsim._p, None, 0.0)
    if r != 0:
        raise RuntimeError(f"detach failed ({r})")
