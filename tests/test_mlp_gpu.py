"""Learned recovery on the GPU: BF16 MFMA products, gradients against the
FP64 reference, training, recovery with the learned Kastaun bracket and
with the learned 1Dvsq start."""
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
    """The learned bracket (model evaluated inside the recovery kernel, gate
    0: every cell) drives the recovery of a 3D torus:
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


@pytest.mark.gpu
def test_in_wave_inference_matches_training_forward(trained):
    """The recovery kernel's inference (one wavefront over 64 states, the
    activations kept in the matrix-core fragments, weights staged in LDS)
    gives the training forward pass's prediction up to its hardware exp and
    reciprocal in the activation (a BF16 rounding flip at most here and
    there)."""
    m, _ = trained
    rng = np.random.default_rng(6)
    n = 1000  # not a multiple of 64: the last wavefront is partial
    feats = np.stack([rng.uniform(-0.5, 3, n), rng.uniform(0, 6, n), rng.uniform(0, 4, n),
                      rng.uniform(0, 6, n), rng.uniform(0, 1, n)], 1).astype(np.float32)
    gw = m.predict_wave(feats)
    gf = m.predict(feats)
    d = np.abs(gw - gf)
    print("wave vs forward: max %.2e mean %.2e" % (d.max(), d.mean()), flush=True)
    assert d.max() <= 2e-3 and d.mean() <= 2e-4


def _cell_metrics(h, cfg):
    """Code-coordinate metric at the interior cell centres of a one-block
    2-D mesh (the oracle's evaluation of the metric the tables hold)."""
    import oracle_py as O
    xg, n = h.block_geom(0)
    s, e = h.s, h.e
    out = {}
    for j in range(s[1], e[1] + 1):
        for i in range(s[0], e[0] + 1):
            x = [0.0, xg[0] + (i + 0.5) * xg[3], xg[1] + (j + 0.5) * xg[4], xg[2] + (s[2] + 0.5) * xg[5]]
            gcov, gcon, _, _ = O.metric_at(3, x, cfg["a"], cfg["h"])
            out[(j, i)] = (gcov, gcon)
    return out


@pytest.mark.gpu
def test_learned_bracket_on_the_2d_torus_workload():
    """The learned bracket on the 2-D MKS torus workload (the 512x256 run at
    half resolution per dimension, Kastaun recovery), trained on seeded
    synthetic draws plus states harvested from the plain run itself:
    - every accepted cell's root mu satisfies |f(mu)| <= 1e-10 (Kastaun's
      master function evaluated by the oracle from the cell's conserved
      state), as every root of the plain run does;
    - P and U stay within the solver tolerance of the plain run after ten
      cycles;
    - on the cells passing the gate (previous recovery at least GATE
      steps; their wavefronts run the model) the learned search takes fewer
      false-position steps."""
    import oracle_py as O
    from pangu_amd import configs, mlp
    from pangu_amd.sim import HipSim
    GATE = 4
    cfg = dict(configs.torus2d(256, 128), recovery=1)
    m = mlp.Mlp(seed=3)
    m.make_dataset(1 << 18, seed=21)
    src = HipSim(cfg)
    src.init()
    harvested = 0
    for _ in range(4):
        src.step(2, "graph")
        harvested += m.harvest(src, 0)
    src.close()
    print("harvested", harvested, flush=True)
    assert harvested > 0.95 * 4 * 256 * 128
    m.train(epochs=8, batch=1024, lr=2e-3)
    runs = {}
    for name in ("plain", "learned"):
        h = HipSim(cfg)
        h.init()
        h.record_roots(True)
        if name == "learned":
            m.attach(h, delta=0.01, gate=GATE)
        h.step(9, "graph")
        prev = h.interior(h.get("c2p_iters")).copy()
        h.step(1, "graph")
        runs[name] = dict(P=h.interior(h.get("prim", 0)), U=h.interior(h.get("cons", 0)),
                          it=h.interior(h.get("c2p_iters")), st=h.interior(h.get("c2p_status")),
                          mu=h.interior(h.get("c2p_root")), prev=prev, h=h)
    geo = _cell_metrics(runs["plain"]["h"], cfg)
    for name, r in runs.items():
        worst, nacc = 0.0, 0
        U, mu, st = r["U"][0], r["mu"][0, 0, 0], r["st"][0, 0, 0]
        for (j, i), (gcov, gcon) in geo.items():
            jj, ii = j - r["h"].s[1], i - r["h"].s[0]
            if st[jj, ii] != 0:
                continue
            nacc += 1
            res = O.kastaun_residual(U[:, 0, jj, ii], mu[jj, ii], cfg["gamma"], gcov, gcon)
            worst = max(worst, res)
        print(name, "accepted", nacc, "of", st.size, "max |f(mu)|", worst, flush=True)
        assert nacc > 0.99 * st.size and worst <= 1e-10
    lp = per_comp_l1(runs["learned"]["P"], runs["plain"]["P"])
    lu = per_comp_l1(runs["learned"]["U"], runs["plain"]["U"])
    print("learned vs plain after 10 cycles: P", max(lp), "U", max(lu), flush=True)
    assert max(lp) <= 1e-8 and max(lu) <= 1e-8
    hard = runs["plain"]["prev"] >= GATE
    ip, il = runs["plain"]["it"][hard].mean(), runs["learned"]["it"][hard].mean()
    print("gated cells %d of %d: steps plain %.2f learned %.2f; all cells plain %.3f learned %.3f" % (
        hard.sum(), hard.size, ip, il, runs["plain"]["it"].mean(), runs["learned"]["it"].mean()), flush=True)
    assert hard.sum() > 0 and il < ip
    for r in runs.values():
        r["h"].close()


@pytest.mark.gpu
def test_learned_1dvsq_start():
    """A model of kind 1 predicts the start of the 1Dvsq Newton iteration
    (log u^2, the iteration starting from v^2 = u^2 / (1 + u^2)), trained on
    synthetic draws plus states harvested from the Orszag-Tang run, whose
    recovery is 1Dvsq (Minkowski):
    - held-out synthetic states: fewer Newton iterations than the start with
      no previous state (v^2 = 0), no state the learned start loses, and
      every learned result within 1e-10 of the 1Dvsq residual's scale;
    - the OT run with the model attached (gate 0, every wavefront): every
      cell recovered in the last stage satisfies the 1Dvsq equation to
      1e-10 of its scale (the oracle's residual from the cell's U, its v^2
      and the entropy the recovery held fixed), as the plain run's cells do,
      and P and U stay within the Newton tolerance of the plain run after
      ten cycles.  The mean iterations against the reference's start (the
      previous primitives' v^2) are reported, not asserted: on a running
      simulation that start is already close."""
    import oracle_py as O
    from pangu_amd import configs, mlp
    from pangu_amd.sim import HipSim
    cfg = configs.orszag_tang(128)
    m = mlp.Mlp(seed=7, kind=1)
    m.make_dataset(1 << 18, seed=41)
    src = HipSim(cfg)
    src.init()
    harvested = 0
    for _ in range(4):
        src.step(2, "graph")
        harvested += m.harvest(src, 0)
    src.close()
    print("harvested", harvested, flush=True)
    assert harvested > 0.95 * 4 * 128 * 128
    loss = m.train(epochs=8, batch=1024, lr=2e-3)
    print("loss first/last", loss[0], loss[-1], flush=True)
    m.make_dataset(1 << 16, seed=42)
    ev = m.evaluate(delta=0.01)
    print("held-out 1Dvsq", ev, flush=True)
    assert ev["failures"] == 0
    assert ev["learned_iters"] < ev["plain_iters"]
    assert ev["max_residual"] <= 1e-10
    runs = {}
    for name in ("plain", "learned"):
        h = HipSim(cfg)
        h.init()
        if name == "learned":
            m.attach(h, delta=0.01, gate=0)
        h.step(10, "graph")
        runs[name] = dict(P=h.interior(h.get("prim", 0)), U=h.interior(h.get("cons", 0)),
                          it=h.interior(h.get("c2p_iters")), st=h.interior(h.get("c2p_status")))
        h.close()
    gam = cfg["gamma"]
    eta = np.diag([-1.0, 1.0, 1.0, 1.0])
    for name, r in runs.items():
        P, U, st = r["P"][0], r["U"][0], r["st"][0, 0, 0]
        worst, nacc = 0.0, 0
        for jj in range(st.shape[0]):
            for ii in range(st.shape[1]):
                if st[jj, ii] != 0:
                    continue
                nacc += 1
                ut = P[2:5, 0, jj, ii]
                usq = float(ut @ ut)
                f, sc = O.vsq_residual(U[:8, 0, jj, ii], P[8, 0, jj, ii], usq / (1.0 + usq), gam, eta, eta)
                worst = max(worst, abs(f) / sc)
        print(name, "accepted", nacc, "of", st.size, "max |f|/scale", worst,
              "mean iterations", float(r["it"].mean()), flush=True)
        assert nacc > 0.99 * st.size and worst <= 1e-10
    lp = per_comp_l1(runs["learned"]["P"], runs["plain"]["P"])
    lu = per_comp_l1(runs["learned"]["U"], runs["plain"]["U"])
    print("learned vs plain after 10 cycles: P", max(lp), "U", max(lu), flush=True)
    assert max(lp) <= 1e-8 and max(lu) <= 1e-8
