"""Pins of the FP64 reference network used to check the GPU training
kernels (CPU): its backward pass against finite differences, BF16 rounding
against known encodings."""
import numpy as np

import mlp_ref as R


def _random_params(seed):
    rng = np.random.default_rng(seed)
    d = {name: rng.normal(0, 0.3, shp) for name, shp in R.SIZES}
    d["mean"] = rng.normal(0, 0.1, 5)
    d["dev"] = rng.uniform(0.5, 2.0, 5)
    return R.pack(d)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _random_params(0)
    feats = rng.normal(0, 1, (16, 5))
    target = rng.normal(0, 1, 16)
    loss, g = R.loss_and_grad(p, feats, target, quantize=False)
    # every parameter except the (frozen) standardisation
    n_train = p.size - 10
    idx = rng.choice(n_train, 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_and_grad(pp, feats, target, False)[0] - R.loss_and_grad(pm, feats, target, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_planted_backward_mistake_is_caught():
    """A transposed weight in the backward pass (a plausible slip) breaks the
    finite-difference agreement."""
    rng = np.random.default_rng(2)
    p = _random_params(3)
    feats = rng.normal(0, 1, (8, 5))
    target = rng.normal(0, 1, 8)
    P = R.unpack(p)
    _, g = R.loss_and_grad(p, feats, target, quantize=False)
    Pw = dict(P)
    Pw["W2"] = P["W2"].T.copy()
    _, g_wrong = R.loss_and_grad(R.pack(Pw), feats, target, quantize=False)
    # gradient of W1 with the transposed W2 differs from the true one
    o = 0
    assert not np.allclose(g[o:o + 320], g_wrong[o:o + 320], rtol=1e-3)


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.14159265, 1e-3, 65504.0], dtype=np.float64)
    got = R.bf16(vals)
    # 1.00390625 = 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125
    assert got[3] == -2.5 and got[4] == 3.140625
    assert abs(got[5] - 1e-3) / 1e-3 < 2 ** -8
    assert got[6] == 65536.0
