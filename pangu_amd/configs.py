"""Standard run configurations.

Each is a plain dict whose keys follow PANGU's input-file parameters
(<parthenon/mesh>, <parthenon/time>, <core>, <metric>, <electron>,
<fm_torus>).  The same dict drives the HIP simulation and, in tests, the
CPU oracle.
"""
import math

_BASE = dict(
    ng=4, dt_init=1e-5, dt_factor=1.3,
    metric=0, a=0.0, h=0.0, gamma=5.0 / 3.0, cfl=0.4,
    riemann=0, limiter=1, recovery=0, enable_B=1, enable_heating=0,
    density_floor=1.0e-6, energy_floor=1.0e-8, density_floor_pow=-1.5, energy_floor_pow=-2.5,
    sigma_max=50.0, lorentz_max=50.0,
    heat_model=0, fel_constant=0.1, gamma_e=4.0 / 3.0, gamma_p=5.0 / 3.0,
    ratio_min=0.001, ratio_max=1000.0, fel_0=0.1,
    limit_kel=1, suppress_highb_heat=0, enforce_positive_dissipation=0,
    excise=0, r_excise=1.0, dexcise=1.0e-8, pexcise=0.333e-12,
    stage_init=0, exchange_after=0, cache_geometry=1,
    problem=1, rin=6.0, rmax=12.0, kappa=1.0e-3, perturbation=4.0e-2, beta=100.0,
    aphi_cut=0.2, seed=0, magnetic_mad=0,
)

PERIODIC, OUTFLOW, REFLECT = 0, 1, 2
METRIC = dict(minkowski=0, bl=1, cks=2, mks=3)
RIEMANN = dict(laxf=0, hll=1)
LIMITER = dict(mc=0, ppm4=1)
RECOVERY = dict(chainer=0, robuster=1, kastaun=1, onedvsq=2)


def orszag_tang(n=256, nb=(1, 1, 1)):
    """2D relativistic Orszag-Tang on [0,1]^2 (n^2 zones), Minkowski, MC,
    LAXF, CT, 1Dvsq recovery, RK2, CFL 0.4."""
    d = dict(_BASE)
    d.update(nx=[n, n, 1], nb=list(nb), xmin=[0.0, 0.0, -0.5], xmax=[1.0, 1.0, 0.5],
             bc=[[PERIODIC] * 2] * 3, metric=0, riemann=0, limiter=0, recovery=2,
             enable_heating=0, problem=1, stage_init=1)
    return d


def _torus_common(d):
    d.update(metric=3, a=0.9375, h=0.7, gamma=5.0 / 3.0, cfl=0.4, enable_B=1,
             enable_heating=1, density_floor=1.0e-5, energy_floor=4.641588833612779e-9,
             problem=2, xmin=[0.2, -1.0, -3.1415926535897], xmax=[6.0, 1.0, 3.1415926535897],
             bc=[[OUTFLOW, OUTFLOW], [REFLECT, REFLECT], [PERIODIC, PERIODIC]])
    return d


def torus2d(nx1=512, nx2=256, nb=(1, 1, 1)):
    """2D FM torus (nx1 x nx2 zones), MKS a=0.9375 h=0.7, PPM4, LAXF,
    chainer, heating, CT, RK2."""
    d = _torus_common(dict(_BASE))
    d.update(nx=[nx1, nx2, 1], nb=list(nb), riemann=0, limiter=1, recovery=0)
    return d


def torus3d(nx=(256, 128, 128), nb=(4, 2, 4)):
    """3D FM torus, by default 256x128x128 zones in 32 meshblocks of
    64x64x32 (the benchmark workload), MKS a=0.9375 h=0.7, PPM4, HLL,
    Kastaun, CT, heating, RK2."""
    d = _torus_common(dict(_BASE))
    d.update(nx=list(nx), nb=list(nb), riemann=1, limiter=1, recovery=1)
    return d


def zones(d):
    return d["nx"][0] * d["nx"][1] * d["nx"][2]


def ensemble(seed=0, n_tori=96, n_boxes_each=121, sizes3d=(16, 32, 64),
             tori_res=((128, 64), (256, 128), (384, 192))):
    """Batched workload: independent MKS tori in one context.

    n_tori 2D Fishbone-Moncrief tori with parameters drawn from a fixed
    seed: a ~ U(0, 0.99), beta ~ logU(30, 300), rin ~ U(5, 7), rmax = rin + 6,
    resolution uniform over tori_res, inner radius 0.9 r_H(a) (inside the
    horizon: x1min = ln(0.9 (1 + sqrt(1 - a^2)))), each with its own
    perturbation seed.  Plus n_boxes_each single-block 3D tori at each of
    the cube sizes in sizes3d (same draws).  Physics as the 3D torus
    configuration: PPM4, HLL, Kastaun, CT, heating.  The defaults give
    ~3.7e6 zones of tori and ~3.6e7 zones of 3D blocks."""
    import random
    rng = random.Random(seed)

    def one(nx, nb):
        a = rng.uniform(0.0, 0.99)
        beta = math.exp(rng.uniform(math.log(30.0), math.log(300.0)))
        rin = rng.uniform(5.0, 7.0)
        rh = 1.0 + math.sqrt(1.0 - a * a)
        d = _torus_common(dict(_BASE))
        d.update(nx=list(nx), nb=list(nb), riemann=1, limiter=1, recovery=1, a=a, beta=beta,
                 rin=rin, rmax=rin + 6.0, seed=rng.randrange(1 << 62))
        d["xmin"] = [math.log(0.9 * rh), -1.0, -3.1415926535897]
        return d

    out = []
    for _ in range(n_tori):
        res = tori_res[rng.randrange(len(tori_res))]
        out.append(one((res[0], res[1], 1), (1, 1, 1)))
    for n in sizes3d:
        for _ in range(n_boxes_each):
            out.append(one((n, n, n), (1, 1, 1)))
    return out


def fm_torus_cks(nx=64, nb=4, levels=4, extent=64.0, enable_B=0, mad=0):
    """The reference's static-refinement run (fm_torus_cks/inputfile): a
    hydrodynamic FM torus (rin 6, rmax 12) around a spinning black hole
    (a = 0.9375) in Cartesian Kerr-Schild coordinates with excision,
    LAXF + PPM4, Kastaun recovery, gamma 4/3, CFL 0.3; nx^3 root cells over
    [-extent, extent]^3 in nb^3 blocks, refined to `levels` nested cubes
    [-extent/2^L, extent/2^L]^3.  enable_B adds the reference's SANE (or,
    with mad, MAD) vector-potential field at beta = 100; the reference input
    runs without it."""
    d = dict(_BASE)
    d.update(nx=[nx] * 3, nb=[nb] * 3, xmin=[-extent] * 3, xmax=[extent] * 3,
             bc=[[OUTFLOW, OUTFLOW]] * 3, dt_init=1e-5, dt_factor=1.3,
             metric=METRIC["cks"], a=0.9375, h=0.7, gamma=4.0 / 3.0, cfl=0.3,
             riemann=RIEMANN["laxf"], limiter=LIMITER["ppm4"], recovery=RECOVERY["robuster"],
             enable_B=enable_B, magnetic_mad=mad, enable_heating=0, density_floor=1e-5,
             energy_floor=1e-7,
             density_floor_pow=-1.5, energy_floor_pow=-2.5,
             excise=1, r_excise=1.0, dexcise=1.0e-8, pexcise=0.333e-10,
             problem=3, rin=6.0, rmax=12.0, kappa=1e-3, perturbation=2e-2)
    d["refine"] = [(L, (-extent / 2 ** L,) * 3, (extent / 2 ** L,) * 3)
                   for L in range(1, levels + 1)]
    return d
