"""Learned Kastaun bracket on the GPU: BF16 MFMA products, gradients against
the FP64 reference, training, and recovery with the learned bracket."""
import numpy as np
import pytest

import mlp_ref as R
from helpers import per_comp_l1, small_configs


@pytest.mark.gpu
def test_mfma_gemm_matches_bf16_product():
    """The 32x32x16 BF16 MFMA tile (operand and accumulator lane layout)
    reproduces the product of the BF16-rounded matrices."""
    from pangu_amd import mlp
    rng = np.random.default_rng(0)
    for M, N, K in ((32, 32, 16), (64, 96, 48), (96, 64, 128)):
        A = rng.normal(0, 1, (M, K)).astype(np.float32)
        B = rng.normal(0, 1, (K, N)).astype(np.float32)
        C = mlp.gemm(A, B)
        ref = R.bf16(A) @ R.bf16(B)
        err = np.abs(C - ref).max() / np.abs(ref).max()
        print(M, N, K, "max rel err", err, flush=True)
        assert err < 1e-5


@pytest.mark.gpu
def test_gradient_matches_fp64_reference():
    from pangu_amd import mlp
    rng = np.random.default_rng(4)
    m = mlp.Mlp(seed=7)
    p = m.params()
    p[-10:-5] = rng.normal(0, 0.2, 5)      # feature mean
    p[-5:] = rng.uniform(0.5, 2.0, 5)      # feature deviation
    p[:-10] += rng.normal(0, 0.05, p.size - 10).astype(np.float32)
    m.set_params(p)
    feats = rng.normal(0, 1, (128, 5)).astype(np.float32)
    target = rng.normal(0, 1, 128).astype(np.float32)
    loss, g = m.gradient(feats, target)
    lref, gref = R.loss_and_grad(m.params().astype(np.float64), feats, target.astype(np.float64))
    assert abs(loss - lref) <= 1e-4 * max(1.0, lref)
    o = 0
    for name, shp in R.SIZES[:-2]:
        n = int(np.prod(shp))
        a, b = g[o:o + n], gref[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(name, "rel err", rel, flush=True)
        assert rel < 2e-2, name
        o += n


@pytest.fixture(scope="module")
def trained():
    from pangu_amd import mlp
    m = mlp.Mlp(seed=1)
    kept = m.make_dataset(1 << 20, seed=2)
    assert kept > 0.9 * (1 << 20)
    loss = m.train(epochs=20, batch=1024, lr=2e-3)
    print("loss per epoch", loss, flush=True)
    m.make_dataset(1 << 17, seed=3)   # held-out states for the evaluation
    return m, loss


@pytest.mark.gpu
def test_training_cuts_kastaun_iterations(trained):
    m, loss = trained
    assert loss[-1] < 0.2 * loss[0]
    ev = m.evaluate(delta=0.01)
    print("evaluation", ev, flush=True)
    assert ev["failures"] == 0
    assert ev["learned_iters"] <= 0.6 * ev["plain_iters"]      # >= 40 % fewer steps
    assert ev["learned_evals"] <= 0.75 * ev["plain_evals"]
    assert ev["max_residual"] <= 1e-9
    # Both paths stop on the same 1e-10 step rule in mu; rho, u and u~ follow
    # from mu through eps = D eod - rho, whose cancellation can amplify the
    # root's last digits (worst case seen ~2e-8 over 1.3e5 states).
    assert ev["max_rel_diff"] <= 1e-6


@pytest.mark.gpu
def test_predict_matches_fp64_reference(trained):
    m, _ = trained
    rng = np.random.default_rng(5)
    feats = np.stack([rng.uniform(0, 3, 256), rng.uniform(0, 6, 256), rng.uniform(0, 4, 256),
                      rng.uniform(0, 6, 256), rng.uniform(0, 1, 256)], 1).astype(np.float32)
    g = m.predict(feats)
    ref = R.forward(m.params().astype(np.float64), feats)
    assert np.abs(g - ref).max() < 1e-3


@pytest.mark.gpu
def test_learned_bracket_in_torus_run(trained):
    """The fused per-cell inference kernel drives the recovery of a 3D torus:
    fields stay within the solver tolerance of the plain run and the
    recovery takes no more false-position steps.  (Most torus cells are cold
    atmosphere with a root near 1, where the plain search already needs only
    about two steps, so the saving here is small.)"""
    from pangu_amd.sim import HipSim
    m, _ = trained
    cfg = small_configs()["torus3d_64x32x32_8blk"]
    runs = []
    for learned in (False, True):
        h = HipSim(cfg)
        h.init()
        if learned:
            m.attach(h, delta=0.01)
        h.step(2, "graph")
        it = h.interior(h.get("c2p_iters"))
        runs.append((h.interior(h.get("prim", 0)), h.interior(h.get("cons", 0)), it.mean()))
        h.close()
    print("mean iterations plain %.2f learned %.2f" % (runs[0][2], runs[1][2]), flush=True)
    assert max(per_comp_l1(runs[1][0], runs[0][0])) <= 1e-9
    assert max(per_comp_l1(runs[1][1], runs[0][1])) <= 1e-9
    assert runs[1][2] <= runs[0][2]
