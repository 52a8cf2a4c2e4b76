"""Run diagnostics (shell fluxes, volume integrals, corner div B).

CPU: the oracle's diagnostics against closed forms for uniform Minkowski
states (fluxes of a boosted, magnetised fluid through x1 = const), a shell
index test across blocks, and planted single-cell divergences.
GPU: pangu_diagnostics against the oracle after a few cycles of several
configurations, bitwise repeatability, and batch members against the same
meshes run alone."""
import numpy as np
import pytest

import oracle_py as O
from pangu_amd import configs

GAM = 5.0 / 3.0


def _uniform_box(nb=(1, 1, 1), nx=(32, 32, 1), xmax=(1.0, 1.0, 0.5)):
    d = configs.orszag_tang(nx[0], nb)
    d.update(nx=list(nx), xmax=list(xmax), gamma=GAM)
    if nx[2] > 1:
        d["xmin"] = [0.0, 0.0, 0.0]
        d["xmax"] = list(xmax)
    return d


def _set_uniform(o, rho, u, ut, B):
    P = o.get("prim", 0)
    P[:, 0] = rho
    P[:, 1] = u
    for a in range(3):
        P[:, 2 + a] = ut[a]
        P[:, 5 + a] = B[a]
    P[:, 8] = 1.0
    o.set("prim", P, 0)
    o.task("conservative", 0)
    return P


def _closed_form(rho, u, ut, B):
    """Minkowski: u^0 = W, u^i = u~^i, b^0 = B.u, b^i = (B^i + b^0 u^i) / W."""
    W = np.sqrt(1.0 + np.dot(ut, ut))
    ucon = np.array([W, *ut])
    eta = np.diag([-1.0, 1.0, 1.0, 1.0])
    ucov = eta @ ucon
    b0 = np.dot(B, ut)
    bcon = np.array([b0, *((np.array(B) + b0 * np.array(ut)) / W)])
    bcov = eta @ bcon
    bsq = bcon @ bcov
    w = rho + GAM * u + bsq
    ptot = (GAM - 1.0) * u + 0.5 * bsq
    T = np.outer(w * ucon, ucov) + ptot * np.eye(4) - np.outer(bcon, bcov)  # T^mu_nu
    return dict(W=W, u1=ucon[1], T10=T[1, 0], T13=T[1, 3], T00=T[0, 0], ucon0=ucon[0])


@pytest.mark.parametrize("nb", [(1, 1, 1), (2, 2, 1)])
def test_oracle_shell_fluxes_uniform_flow(nb):
    rho, u, ut, B = 2.0, 0.5, (0.3, -0.1, 0.0), (0.7, 0.2, 0.0)
    d = _uniform_box(nb)
    o = O.OracleSim(d)
    o.init()
    _set_uniform(o, rho, u, ut, B)
    cf = _closed_form(rho, u, ut, B)
    A = (d["xmax"][1] - d["xmin"][1]) * (d["xmax"][2] - d["xmin"][2])
    V = A * (d["xmax"][0] - d["xmin"][0])
    for shell in (0, 5, 31):
        g = o.diagnostics(shell)
        np.testing.assert_allclose(g["mdot"], -rho * cf["u1"] * A, rtol=1e-12)
        np.testing.assert_allclose(g["edot"], cf["T10"] * A, rtol=1e-12)
        assert abs(g["jdot"]) <= 1e-15
        np.testing.assert_allclose(g["phi"], 0.5 * B[0] * A, rtol=1e-12)
        np.testing.assert_allclose(g["mass"], rho * cf["W"] * V, rtol=1e-12)
        np.testing.assert_allclose(g["energy"], (cf["T00"] + rho * cf["ucon0"]) * V, rtol=1e-12)
        assert g["divb_max"] == 0.0
        assert g["ub_max"] == 0.7


def test_oracle_shell_index_across_blocks():
    """rho = 1 + (global x1 index): the shell picks the cells of its own
    global index, also in the second block along x1."""
    d = _uniform_box((2, 1, 1))
    o = O.OracleSim(d)
    o.init()
    P = _set_uniform(o, 1.0, 0.5, (0.3, 0.0, 0.0), (0.0, 0.0, 0.0))
    nxb = d["nx"][0] // 2
    for b in range(2):
        xd, n = o.block_geom(b)
        s0, e0 = n[3], n[6]
        for i in range(s0, e0 + 1):
            P[b, 0, :, :, i] = 1.0 + b * nxb + (i - s0)
    o.set("prim", P, 0)
    A = (d["xmax"][1] - d["xmin"][1]) * (d["xmax"][2] - d["xmin"][2])
    for shell in (3, nxb - 1, nxb, 2 * nxb - 2):
        g = o.diagnostics(shell)
        np.testing.assert_allclose(g["mdot"], -(1.0 + shell) * 0.3 * A, rtol=1e-12)


@pytest.mark.parametrize("nx", [(16, 16, 1), (8, 8, 8)])
def test_oracle_planted_divergence(nx):
    """One perturbed U_B1 cell gives |delta| / (2 dx1) at its corners in 2-D
    and |delta| / (4 dx1) in 3-D."""
    d = _uniform_box((1, 1, 1), nx, xmax=(1.0, 1.0, 1.0 if nx[2] > 1 else 0.5))
    o = O.OracleSim(d)
    o.init()
    _set_uniform(o, 1.0, 0.5, (0.0, 0.0, 0.0), (0.3, 0.1, 0.05 if nx[2] > 1 else 0.0))
    U = o.get("cons", 0)
    xd, n = o.block_geom(0)
    s = n[3:6]
    k = s[2] + (nx[2] // 2 if nx[2] > 1 else 0)
    U[0, 5, k, s[1] + 4, s[0] + 4] += 1e-3
    o.set("cons", U, 0)
    g = o.diagnostics(0)
    dx1 = 1.0 / nx[0]
    want = 1e-3 / ((4 if nx[2] > 1 else 2) * dx1)
    np.testing.assert_allclose(g["divb_max"], want, rtol=1e-9)


def _cases():
    return {
        "torus2d": configs.torus2d(64, 32),
        "torus3d": configs.torus3d((32, 16, 16), (2, 2, 2)),
        "ot": configs.orszag_tang(32, (2, 1, 1)),
    }


def _compare(h, o, shells, label):
    for sh in shells:
        gh = h.diagnostics(sh)
        go, mag = o.diagnostics(sh, magnitudes=True)
        print(label, sh, {k: (gh[k], go[k]) for k in gh}, flush=True)
        for k in ("mdot", "edot", "jdot", "phi", "mass", "energy"):
            # sums of signed terms: rounding scales with the sum of |terms|
            assert abs(gh[k] - go[k]) <= 1e-12 * mag[k] + 1e-300, (k, gh[k], go[k], mag[k])
        assert abs(gh["ub_max"] - go["ub_max"]) <= 1e-13 * max(go["ub_max"], 1e-300)
        assert abs(gh["divb_max"] - go["divb_max"]) <= 1e-13 * max(go["ub_max"], 1e-300) * 64


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["torus2d", "torus3d", "ot"])
def test_diagnostics_match_oracle(name):
    from pangu_amd.sim import HipSim
    d = _cases()[name]
    o = O.OracleSim(d)
    o.init()
    o.step(2)
    h = HipSim(d)
    h.init()
    h.step(2, "graph")
    nx1 = d["nx"][0]
    _compare(h, o, (0, nx1 // 3, nx1 - 1), name)
    h.close()


@pytest.mark.gpu
def test_diagnostics_repeat_bitwise_and_reject_bad_shell():
    from pangu_amd.sim import HipSim
    d = configs.torus3d((32, 16, 16), (2, 2, 2))
    h = HipSim(d)
    h.init()
    h.step(1, "graph")
    a = h.diagnostics(7)
    b = h.diagnostics(7)
    assert all(a[k] == b[k] for k in a)
    with pytest.raises(RuntimeError):
        h.diagnostics(32)
    h.close()


@pytest.mark.gpu
def test_batch_diagnostics_equal_each_mesh_alone():
    from pangu_amd.sim import HipSim
    d2 = configs.torus2d(64, 32)
    d2.update(riemann=1, recovery=1)  # a batch shares one set of physics options
    cfgs = [d2, configs.torus3d((32, 16, 16), (2, 2, 2))]
    shells = [10, 5]
    hb = HipSim(cfgs)
    hb.init()
    hb.step(2, "graph")
    gb = hb.diagnostics(shells)
    hb.close()
    for m, d in enumerate(cfgs):
        h = HipSim(d)
        h.init()
        h.step(2, "graph")
        g = h.diagnostics(shells[m])
        h.close()
        assert all(g[k] == gb[m][k] for k in g), (m, g, gb[m])


def _cks_refined():
    return configs.fm_torus_cks(nx=32, nb=4, levels=2, extent=32.0)


def test_oracle_volume_integrals_skip_excised_cells():
    """The reference's refined CKS torus after one cycle: U is non-finite in
    excised cells (r < r_excise, never recovered), yet mass, energy, |U_B|
    and div B stay finite because those cells are left out; shell integrals
    are skipped with shell -1."""
    d = _cks_refined()
    o = O.OracleSim(d)
    o.init()
    o.step(1)
    U = o.interior(o.get("cons", 0))
    assert not np.isfinite(U).all()  # the cells the rule is for exist
    g = o.diagnostics(-1)
    for k in ("mass", "energy", "ub_max", "divb_max"):
        assert np.isfinite(g[k]), (k, g[k])
    assert g["mass"] > 0.0 and g["mdot"] == 0.0 and g["phi"] == 0.0


@pytest.mark.gpu
def test_diagnostics_refined_cks_finite_and_match_oracle():
    from pangu_amd.sim import HipSim
    d = _cks_refined()
    o = O.OracleSim(d)
    o.init()
    o.step(1)
    h = HipSim(d)
    h.init()
    h.step(1, "graph")
    gh = h.diagnostics(-1)
    go, mag = o.diagnostics(-1, magnitudes=True)
    print("refined cks", gh, go, flush=True)
    for k in ("mass", "energy"):
        assert np.isfinite(gh[k]) and abs(gh[k] - go[k]) <= 1e-12 * mag[k], (k, gh[k], go[k])
    with pytest.raises(RuntimeError):
        h.diagnostics(3)  # shell integrals need a single-level mesh
    h.close()
