"""The oracle's operation-counting build (oracle/counting.h), which measures
the reference algorithm's operations per zone.

Pinned by hand counts of two tasks whose arithmetic is fixed by their
formulas (flux divergence and update), and by bitwise equality of the
counted run with the ordinary oracle (counting must not change a value)."""
import numpy as np

import oracle_py as O
from pangu_amd import configs


def _per_zone(counts, task, cfg, cycles=1):
    zones = cfg["nx"][0] * cfg["nx"][1] * cfg["nx"][2] * cycles
    row = counts[O.TASK_NAMES.index(task)]
    return {op: row[i] / zones for i, op in enumerate(O.OP_NAMES)}


def test_flux_divergence_and_update_hand_counts_3d():
    """Per component and stage: 3 directions x (2 mul + 1 sub), 2 adds to
    combine them, one division by the volume; the update is one mul and one
    add.  Ten components, two stages."""
    d = configs.torus3d((16, 8, 8), (1, 1, 1))
    c = O.op_counts(d, 1)
    fd = _per_zone(c, "flux_div", d)
    assert fd["mul"] == 2 * 10 * 6 and fd["add"] == 2 * 10 * 5 and fd["div"] == 2 * 10
    up = _per_zone(c, "update", d)
    assert up["mul"] == 2 * 10 and up["add"] == 2 * 10 and up["div"] == 0


def test_flux_divergence_hand_counts_2d():
    """Orszag-Tang: two directions, nine components."""
    d = configs.orszag_tang(16)
    c = O.op_counts(d, 1)
    fd = _per_zone(c, "flux_div", d)
    assert fd["mul"] == 2 * 9 * 4 and fd["add"] == 2 * 9 * 3 and fd["div"] == 2 * 9


def test_counting_changes_no_value():
    d = dict(configs.torus3d((16, 8, 8), (2, 1, 1)), cache_geometry=0)
    counts, prim = O.op_counts(d, 2, return_prim=True)
    o = O.OracleSim(d, nthreads=1)
    o.init()
    o.step(2)
    ref = o.get("prim", 0).reshape(-1)
    np.testing.assert_array_equal(prim, ref)
    # the tabulated-geometry run does far fewer metric evaluations
    d2 = dict(d, cache_geometry=1)
    c2 = O.op_counts(d2, 2)
    assert c2[:, O.OP_NAMES.index("sincos")].sum() < 0.01 * counts[:, O.OP_NAMES.index("sincos")].sum()
