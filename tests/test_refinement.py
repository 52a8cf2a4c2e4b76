"""Static mesh refinement in the oracle (CPU): the block tree, ghost cells
across levels and flux correction (oracle/refine.cpp).

Parthenon, which the reference relies on for these, is an empty submodule in
the reference checkout, so no reference output exists for them.  They are
pinned instead by hand counts of the block tree, by exactness on linear
data, by bitwise equality of an everywhere-refined mesh with the uniform mesh
of the same resolution, and by mass conservation across refinement faces;
planted mistakes show that the pins catch a wrong slope factor and a missing
flux correction."""
import collections

import numpy as np
import pytest

import oracle_py as O
from pangu_amd import configs


def cks_smr_mesh():
    """The block tree of fm_torus_cks/inputfile: 64^3 cells over [-64, 64]^3
    in 16^3 blocks, static refinement regions [-64/2^L, 64/2^L]^3 at levels
    L = 1..4 (inputfile:87-122)."""
    d = dict(configs._BASE)
    d.update(nx=[64, 64, 64], nb=[4, 4, 4], xmin=[-64.0] * 3, xmax=[64.0] * 3,
             bc=[[configs.OUTFLOW] * 2] * 3, metric=2, a=0.9375, h=0.7, problem=1,
             enable_B=0, recovery=1, limiter=1, riemann=0)
    d["refine"] = [(L, (-64.0 / 2 ** L,) * 3, (64.0 / 2 ** L,) * 3) for L in (1, 2, 3, 4)]
    return d


def _locs(o):
    return [o.block_loc(b) for b in range(o.nblocks)]


def test_fm_torus_cks_block_tree():
    """Hand count: each level-L region covers the middle 2x2x2 blocks of
    level L-1 (edges of level L-1 blocks fall exactly on +-64/2^L), so level L
    replaces 8 blocks of level L-1 by 64: 64-8 = 56 leaves on levels 0-3 and
    64 on level 4, 288 in all.  Neighbours already differ by one level, so
    the 2:1 rule adds nothing."""
    o = O.OracleSim(cks_smr_mesh(), nthreads=1)
    c = collections.Counter(l[0] for l in _locs(o))
    assert o.nblocks == 288
    assert dict(c) == {0: 56, 1: 56, 2: 56, 3: 56, 4: 64}
    # level-4 leaves are exactly the 4x4x4 blocks of level-4 locations 30..33
    l4 = sorted(l[1:] for l in _locs(o) if l[0] == 4)
    assert l4 == sorted((i, j, k) for k in range(30, 34) for j in range(30, 34) for i in range(30, 34))


def test_two_to_one_rule_refines_neighbours():
    """2-D, 4x4 root blocks on [0,1]^2; one level-2 region [0.26, 0.3]^2.
    Level-2 blocks are 1/16 wide: the region spans block 4 (upper edge 0.3125
    is the first above 0.26 and the first >= 0.3), widened to the pair 4-5.
    Creating them refines level-1 block (2,2), whose level-1 neighbours
    (1..3, 1..3) lie in root blocks (0,0), (1,0), (0,1) and (1,1): all four
    roots are refined.  Leaves: 12 roots, 4x4 - 1 = 15 level-1, 4 level-2."""
    d = configs.orszag_tang(64, (4, 4, 1))
    d["refine"] = [(2, (0.26, 0.26, -0.5), (0.3, 0.3, 0.5))]
    o = O.OracleSim(d, nthreads=1)
    locs = _locs(o)
    c = collections.Counter(l[0] for l in locs)
    assert dict(c) == {0: 12, 1: 15, 2: 4}
    assert sorted(l[1:3] for l in locs if l[0] == 2) == [(4, 4), (4, 5), (5, 4), (5, 5)]
    roots = {l[1:3] for l in locs if l[0] == 0}
    assert not roots & {(0, 0), (1, 0), (0, 1), (1, 1)}


def _set_linear(o, coef, container=1):
    """P_q = c0 + c1 x + c2 y (+ c3 z) at every interior cell centre, zeros
    in ghost cells; returns the exact values for every cell."""
    P = o.get("prim", container)
    exact = np.zeros_like(P)
    for b in range(o.nblocks):
        xd, n = o.block_geom(b)
        xs = [xd[d] + (np.arange(n[d]) + 0.5) * xd[3 + d] for d in range(3)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        f = coef[0] + coef[1] * X + coef[2] * Y + (coef[3] * Z if n[2] > 1 else 0.0)
        for q in range(P.shape[1]):
            exact[b, q] = f * (q + 1)
    P[...] = 0.0
    s, e = o.s, o.e
    P[:, :, s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1] = exact[:, :, s[2]:e[2] + 1, s[1]:e[1] + 1,
                                                                 s[0]:e[0] + 1]
    o.set("prim", P, container)
    return exact


def _ghost_error(o, exact, container=1):
    """Largest |ghost - exact| over ghost cells inside the periodic domain."""
    P = o.get("prim", container)
    return np.abs(P - exact).max()


@pytest.mark.parametrize("plant", ["none", "prolong_slope"])
def test_linear_data_is_exchanged_exactly(plant):
    """Linear data: restriction of 2x2 fine cells is the value at the coarse
    centre, and minmod slopes equal the exact differences, so +-1/4 of them
    gives the fine centres.  Dyadic coefficients and cell centres keep every
    operation exact; a domain of [0, 1]^2 with periodic wrap is not linear
    across the wrap, so the data is placed on a box with outflow faces."""
    d = configs.orszag_tang(64, (4, 4, 1))
    d["refine"] = [(1, (0.3, 0.3, -0.5), (0.7, 0.7, 0.5))]
    d["bc"] = [[configs.OUTFLOW] * 2, [configs.OUTFLOW] * 2, [configs.PERIODIC] * 2]
    o = O.OracleSim(d, nthreads=1)
    O.plant(plant)
    try:
        exact = _set_linear(o, (0.5, 0.25, -0.125, 0.0))
        o.task("exchange", 0, 1)
    finally:
        O.plant("none")
    P = o.get("prim", 1)
    # compare ghost cells that lie inside the domain (physical ones are 0)
    err = 0.0
    for b in range(o.nblocks):
        xd, n = o.block_geom(b)
        xs = [xd[dd] + (np.arange(n[dd]) + 0.5) * xd[3 + dd] for dd in range(2)]
        Y, X = np.meshgrid(xs[1], xs[0], indexing="ij")
        inside = (X > 0) & (X < 1) & (Y > 0) & (Y < 1)
        err = max(err, np.abs(P[b, :, 0][:, inside] - exact[b, :, 0][:, inside]).max())
    if plant == "none":
        assert err == 0.0
    else:
        assert err > 1e-4


def test_everywhere_refined_equals_uniform_mesh():
    """A region covering the whole domain at level 1 gives the blocks (and
    block edges) of the uniform mesh with twice the blocks per dim; every
    ghost is then a same-level copy, so one cycle matches bit for bit."""
    base = configs.orszag_tang(32, (2, 2, 1))
    base["refine"] = [(1, (0.0, 0.0, -0.5), (1.0, 1.0, 0.5))]
    fine = configs.orszag_tang(32, (4, 4, 1))
    fine["nx"] = [64, 64, 1]
    a, b = O.OracleSim(base, nthreads=1), O.OracleSim(fine, nthreads=1)
    assert a.nblocks == b.nblocks == 16
    a.init()
    b.init()
    a.step(2)
    b.step(2)
    la, lb = _locs(a), _locs(b)
    Pa, Pb = a.get("prim", 0), b.get("prim", 0)
    for i, l in enumerate(la):
        j = lb.index((0,) + l[1:])
        np.testing.assert_array_equal(Pa[i], Pb[j])
    assert a.dt == b.dt


def _mass(o, c=0):
    U = o.get("cons", c)
    tot = 0.0
    for b in range(o.nblocks):
        xd, n = o.block_geom(b)
        s, e = n[3:6], n[6:9]
        tot += U[b, 0, s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1].sum() * xd[3] * xd[4]
    return tot


@pytest.mark.parametrize("plant", ["none", "no_flux_correction"])
def test_flux_correction_conserves_mass_across_levels(plant):
    """Periodic OT with a refined centre, ghosts exchanged after recovery:
    with flux correction each coarse face flux equals the mean of the fine
    ones, so mass leaves one level exactly as it enters the other (drift at
    round-off); without it the drift is many orders larger."""
    d = configs.orszag_tang(64, (4, 4, 1))
    d["refine"] = [(1, (0.3, 0.3, -0.5), (0.7, 0.7, 0.5))]
    d["exchange_after"] = 1
    o = O.OracleSim(d, nthreads=4)
    o.init()
    o.step(1)
    m0 = _mass(o)
    O.plant(plant)
    try:
        o.step(6)
    finally:
        O.plant("none")
    drift = abs(_mass(o) - m0) / m0
    if plant == "none":
        assert drift < 1e-14
    else:
        assert drift > 1e-9


def _specific_angular_momentum(o, kind, a, h, cks, rho_min, nsample=400):
    """Fishbone-Moncrief's l = u^t u_phi of torus cells (rho > rho_min) from
    the initial primitives: u from the state at the cell centre, u_phi the
    KS azimuthal component (x3 in MKS; -y u_x + x u_y in CKS, since
    dx/dphi = -y and dy/dphi = x at fixed KS r, theta)."""
    P = o.get("prim", 0)
    out = []
    rng = np.random.default_rng(0)
    cand = []
    for b in range(o.nblocks):
        xd, n = o.block_geom(b)
        s, e = n[3:6], n[6:9]
        rho = P[b, 0]
        for k, j, i in np.argwhere(rho[s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1] > rho_min):
            cand.append((b, k + s[2], j + s[1], i + s[0], xd))
    for idx in rng.choice(len(cand), size=min(nsample, len(cand)), replace=False):
        b, k, j, i, xd = cand[idx]
        x = [0.0, xd[0] + (i + 0.5) * xd[3], xd[1] + (j + 0.5) * xd[4], xd[2] + (k + 0.5) * xd[5]]
        gcov, gcon, _, ok = O.metric_at(kind, x, a, h)
        p = np.zeros(10)
        p[0:5] = P[b, 0:5, k, j, i]
        st = O.state(p, gcov, gcon)
        uc = st["ucov"]
        uphi = (-x[2] * uc[1] + x[1] * uc[2]) if cks else uc[3]
        out.append(st["ucon"][0] * uphi)
    return np.array(out)


@pytest.mark.parametrize("plant", ["none", "cks_jacobian"])
def test_cks_torus_has_the_mks_torus_angular_momentum(plant):
    """The FM torus has one angular momentum l = u^t u_phi in every cell
    (Fishbone & Moncrief 1976; BL, KS and CKS share u^t and u_phi when
    u^r = 0, the transformations being t_KS = t_BL + f(r),
    phi_KS = phi_BL + g(r) and the CKS map at fixed r, theta).  The CKS initial data (fm_torus_cks/problem_generator.cpp)
    must therefore give, cell by cell, the l of the MKS torus initial data
    (fm_torus_mks), which the equilibrium pins of the MKS path cover; a
    wrong Jacobian sign breaks it."""
    mks = configs.torus2d(64, 32, (1, 1, 1))
    om = O.OracleSim(mks, nthreads=4)
    om.init()
    l_mks = _specific_angular_momentum(om, 3, mks["a"], mks["h"], False, 0.05)
    assert l_mks.std() < 1e-10 * abs(l_mks.mean())
    cks = configs.fm_torus_cks(nx=32, nb=4, levels=2, extent=32.0)
    O.plant(plant)
    try:
        oc = O.OracleSim(cks, nthreads=4)
        oc.init()
    finally:
        O.plant("none")
    l_cks = _specific_angular_momentum(oc, 2, cks["a"], cks["h"], True, 0.05)
    err = np.abs(l_cks - l_mks.mean()).max() / abs(l_mks.mean())
    if plant == "none":
        assert err < 1e-10
    else:
        assert err > 1e-3


def _cks_mhd_torus(mad=0):
    return configs.fm_torus_cks(nx=32, nb=2, levels=0, extent=32.0, enable_B=1, mad=mad)


@pytest.mark.parametrize("mad", [0, 1])
@pytest.mark.parametrize("plant", ["none", "cks_curl"])
def test_cks_torus_field_is_divergence_free(plant, mad):
    """B = curl A / sqrt|g| by central differences (problem_generator.cpp:
    548-578): the central-difference divergence of sqrt|g| B is then zero
    identically (central differences commute), to round-off, over interior
    cells whose stencil stays in their block, whatever the potential (SANE,
    or MAD with its extra radial factor).  A wrong curl sign breaks it."""
    cfg = _cks_mhd_torus(mad)
    O.plant(plant)
    try:
        o = O.OracleSim(cfg, nthreads=4)
        o.init()
    finally:
        O.plant("none")
    P = o.get("prim", 0)
    iB = 5
    worst, scale = 0.0, 0.0
    for b in range(o.nblocks):
        xd, n = o.block_geom(b)
        s, e = n[3:6], n[6:9]
        dx = xd[3:6]
        X = [xd[d] + (np.arange(n[d]) + 0.5) * xd[3 + d] for d in range(3)]
        sg = np.zeros(tuple(int(v) for v in n[2::-1][:3]))
        # sqrt|g| at the centres the divergence reads (interior of the block)
        sg[...] = np.nan
        for k in range(s[2], e[2] + 1):
            for j in range(s[1], e[1] + 1):
                for i in range(s[0], e[0] + 1):
                    _, _, gdet, _ = O.metric_at(2, [0.0, X[0][i], X[1][j], X[2][k]], cfg["a"], 0.0)
                    sg[k, j, i] = np.sqrt(abs(gdet))
        F = [P[b, iB + d] * sg for d in range(3)]
        sl = (slice(s[2] + 1, e[2]), slice(s[1] + 1, e[1]), slice(s[0] + 1, e[0]))
        div = np.zeros_like(F[0][sl])
        for d, ax in enumerate((2, 1, 0)):
            hi = [slice(a.start, a.stop) for a in sl]
            lo = [slice(a.start, a.stop) for a in sl]
            hi[ax] = slice(sl[ax].start + 1, sl[ax].stop + 1)
            lo[ax] = slice(sl[ax].start - 1, sl[ax].stop - 1)
            div += (F[d][tuple(hi)] - F[d][tuple(lo)]) / (2.0 * dx[d])
        worst = max(worst, np.abs(div).max())
        scale = max(scale, max(np.abs(F[d][sl]).max() for d in range(3)) / min(dx))
    assert scale > 0
    if plant == "none":
        assert worst <= 1e-12 * scale
    else:
        assert worst > 1e-3 * scale


@pytest.mark.parametrize("mad", [0, 1])
def test_cks_torus_field_meets_beta(mad):
    """After the normalisation (:712-730) the largest b^2 over the interior
    is 2 (gamma - 1) (u_max / rho_max) / beta, with the maxima over entire
    blocks (:653-704), for the SANE and the MAD potential alike."""
    cfg = _cks_mhd_torus(mad)
    o = O.OracleSim(cfg, nthreads=4)
    o.init()
    P = o.get("prim", 0)
    umax, rhomax = P[:, 1].max(), P[:, 0].max()
    bsq = 0.0
    for b in range(o.nblocks):
        xd, n = o.block_geom(b)
        s, e = n[3:6], n[6:9]
        for k, j, i in np.argwhere(np.abs(P[b, 5:8, s[2]:e[2] + 1, s[1]:e[1] + 1, s[0]:e[0] + 1]).sum(0) > 0):
            k, j, i = k + s[2], j + s[1], i + s[0]
            x = [0.0] + [xd[d] + ((i, j, k)[d] + 0.5) * xd[3 + d] for d in range(3)]
            gcov, gcon, _, _ = O.metric_at(2, x, cfg["a"], 0.0)
            p = np.zeros(10)
            p[0:8] = P[b, 0:8, k, j, i]
            bsq = max(bsq, O.state(p, gcov, gcon)["bsq"])
    want = 2.0 * (cfg["gamma"] - 1.0) * (umax / rhomax) / cfg["beta"]
    assert bsq > 0 and abs(bsq / want - 1.0) < 1e-12
