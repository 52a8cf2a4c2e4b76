"""HIP path vs the CPU oracle on identical initial data (GPU)."""
import numpy as np
import pytest

import oracle_py as O
from helpers import per_comp_l1, small_configs

CASES = small_configs()


def _pair(cfg):
    from pangu_amd.sim import HipSim
    o = O.OracleSim(cfg)
    o.init()
    h = HipSim(cfg)
    h.init()
    return o, h


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_initial_data_identical(name):
    o, h = _pair(CASES[name])
    assert np.array_equal(o.get("prim", 0), h.get("prim", 0))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["tasks", "fused", "graph"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_one_step_parity(name, mode):
    cfg = CASES[name]
    o, h = _pair(cfg)
    o.step(1)
    h.step(1, mode)
    for c in (0, 1):
        lp = per_comp_l1(h.interior(h.get("prim", c)), o.interior(o.get("prim", c)))
        lu = per_comp_l1(h.interior(h.get("cons", c)), o.interior(o.get("cons", c)))
        print(name, mode, "container", c, "P", max(lp), "U", max(lu))
        assert max(lp) <= 1e-12, (c, lp)
        assert max(lu) <= 1e-12, (c, lu)
    # recovery flags (after recovery, before FixRecovery), status codes and
    # floor masks match bit for bit
    shp = (o.nblocks, 1, o.n[2], o.n[1], o.n[0])
    for f in ("flag_recovery", "c2p_status", "c2p_iters"):
        a = o.interior(o.get(f).reshape(shp))
        b = h.interior(h.get(f).reshape(shp))
        print(f, "mismatches", int((a != b).sum()))
        assert np.array_equal(a, b), f
    assert np.array_equal(o.get("mask").reshape(-1), h.get("mask").reshape(-1))
    assert abs(o.dt - h.dt) <= 1e-15 * o.dt
