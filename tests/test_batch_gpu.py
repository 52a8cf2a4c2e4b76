"""Batched ensemble form (GPU): many independent meshes of different shapes
and spacetimes in one context must give, mesh by mesh, exactly what each
mesh gives when run alone, and match the oracle."""
import numpy as np
import pytest

import oracle_py as O
from helpers import per_comp_l1
from pangu_amd import configs


def _small_ensemble():
    return configs.ensemble(seed=3, n_tori=4, n_boxes_each=2, sizes3d=(8, 16),
                            tori_res=((64, 32), (96, 48)))


def _member_blocks(batch, m, field, container=0):
    """Per-block arrays of mesh m from a batched run."""
    _, _, first, nb = batch.mesh_state(m)
    flat = batch.get(field, container)
    blocks = batch.blocks(flat) if not batch.uniform else list(flat)
    return [blocks[first + i] for i in range(nb)], first


@pytest.mark.gpu
def test_batch_equals_each_mesh_alone():
    from pangu_amd.sim import HipSim
    ens = _small_ensemble()
    batch = HipSim(ens)
    batch.init()
    batch.step(3, "graph")
    for m, cfg in enumerate(ens):
        one = HipSim(cfg)
        one.init()
        one.step(3, "graph")
        dt, tm, first, nb = batch.mesh_state(m)
        assert dt == one.dt and tm == one.time, m
        for fld in ("prim", "cons"):
            for c in (0, 1):
                mine, _ = _member_blocks(batch, m, fld, c)
                ref = one.get(fld, c)
                for i in range(nb):
                    assert np.array_equal(mine[i], ref[i]), (m, fld, c, i)
        one.close()
    batch.close()


@pytest.mark.gpu
@pytest.mark.parametrize("m", [0, 5])
def test_batch_member_matches_oracle(m):
    from pangu_amd.sim import HipSim
    ens = _small_ensemble()
    batch = HipSim(ens)
    batch.init()
    batch.step(1, "graph")
    o = O.OracleSim(ens[m])
    o.init()
    o.step(1)
    for fld in ("prim", "cons"):
        mine, first = _member_blocks(batch, m, fld)
        ref = o.get(fld, 0)
        for i, blk in enumerate(mine):
            l = per_comp_l1(batch.block_interior(blk, first + i)[None],
                            o.interior(ref[i:i + 1]))
            print("mesh", m, fld, "block", i, max(l), flush=True)
            assert max(l) <= 1e-12
    dt, tm, _, _ = batch.mesh_state(m)
    assert abs(dt - o.dt) <= 1e-15 * o.dt


@pytest.mark.gpu
def test_autotune_restores_state_and_keeps_results():
    from pangu_amd.sim import HipSim
    ens = _small_ensemble()
    runs = []
    for tune in (False, True):
        h = HipSim(ens)
        h.init()
        h.step(1, "graph")
        if tune:
            choice = h.autotune(cycles=2, use_cache=False)
            print("autotune", choice, flush=True)
        h.step(2, "graph")
        runs.append((h.get("prim", 0), h.get("cons", 0), [h.mesh_state(m)[:2] for m in range(len(ens))]))
        h.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


@pytest.mark.gpu
def test_plain_loop_equals_batch():
    from pangu_amd.sim import HipSim, plain_loop
    ens = _small_ensemble()
    batch = HipSim(ens)
    batch.init()
    batch.step(2, "graph")
    singles = [HipSim(c) for c in ens]
    for s in singles:
        s.init()
    ms = plain_loop(singles, 2, "graph")
    assert ms > 0
    for m, s in enumerate(singles):
        mine, _ = _member_blocks(batch, m, "prim")
        ref = s.get("prim", 0)
        for i, blk in enumerate(mine):
            assert np.array_equal(blk, ref[i]), m
        assert batch.mesh_state(m)[:2] == (s.dt, s.time)


@pytest.mark.gpu
def test_full_ensemble_members_equal_single_runs():
    """The benchmark's own ensemble (96 2-D tori and 363 single-block 3-D
    tori, 459 meshes, about 4e7 zones): two cycles of the batched context
    with the autotuner's launch choice, then three members (the first
    torus, a 32^3 box and the last 64^3 box) each run alone: fields, dt and
    time bit for bit."""
    from pangu_amd.sim import HipSim
    ens = configs.ensemble(seed=0)
    assert len(ens) == 459
    batch = HipSim(ens)
    batch.init()
    print("autotune", batch.autotune(cycles=1), flush=True)
    batch.step(2, "graph")
    # member 224 is a 32^3 torus drawn inside the marginally bound orbit of
    # its spin (a = 0.002, r_in = 5.2): no torus, the mesh holds atmosphere
    members = [0, 96 + 121 + 7, len(ens) - 1]
    singles = {}
    for m in members:
        one = HipSim(ens[m])
        one.init()
        one.step(2, "graph")
        dt, tm, _, _ = batch.mesh_state(m)
        print("member", m, ens[m]["nx"], "dt", dt, one.dt, flush=True)
        assert dt == one.dt and tm == one.time, m
        singles[m] = one
    for fld in ("prim", "cons"):
        for c in (0, 1):
            blocks = batch.blocks(batch.get(fld, c))  # one copy of the arena per field
            if fld == "prim":
                for m in range(len(ens)):
                    _, _, first, nb = batch.mesh_state(m)
                    for i in range(nb):
                        assert np.isfinite(batch.block_interior(blocks[first + i], first + i)).all(), m
            for m in members:
                _, _, first, nb = batch.mesh_state(m)
                ref = singles[m].get(fld, c)
                for i in range(nb):
                    assert np.array_equal(blocks[first + i], ref[i]), (m, fld, c, i)
            del blocks
    for one in singles.values():
        one.close()
    batch.close()
