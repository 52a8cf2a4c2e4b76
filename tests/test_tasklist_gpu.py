"""The task-level C ABI driven from Python on caller-owned buffers laid out
per (meshblock, variable), as a Parthenon MeshBlockPack holds them
(pangu_amd.tasklist over the ctypes mirror pangu_amd.abi), against the
package's own driver in its task mode (one arena per field): the same
entry points in the same order must give the same bits (GPU)."""
import numpy as np
import pytest

from helpers import small_configs
from pangu_amd import configs

CASES = dict(small_configs())
CASES["torus3d_chainer_laxf_mc"] = dict(configs.torus3d((32, 16, 16), (2, 2, 2)), recovery=0,
                                        riemann=0, limiter=0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_tasklist_equals_package_driver(name):
    from pangu_amd.sim import HipSim
    from pangu_amd.tasklist import TaskList
    cfg = CASES[name]
    print("case", name, flush=True)
    h = HipSim(cfg)
    h.init()
    t = TaskList(cfg)
    t.set("prim", h.get("prim", 0), "base")
    for n in (1, 2):
        h.step(1, "tasks")
        t.cycle(1)
        for fld in ("prim", "cons"):
            for c, cname in ((0, "base"), (1, "1")):
                a, b = t.get(fld, cname), h.get(fld, c)
                same = np.array_equal(t.interior(a), h.interior(b))
                print(name, "cycle", n, fld, cname, "bit-equal" if same else "DIFFERENT", flush=True)
                assert same, (name, n, fld, cname)
        assert t.dt == h.dt and t.time == h.time
    t.close()
    h.close()
