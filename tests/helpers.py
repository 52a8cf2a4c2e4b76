"""Shared test helpers: run configurations and comparison metrics."""
import numpy as np

from pangu_amd import configs


def rel_l1(a, b):
    """Relative L1 sum|a-b| / sum|b| (interior arrays); 0 when both vanish."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).sum()
    num = np.abs(a - b).sum()
    if den == 0.0:
        return float(num)
    return float(num / den)


def per_comp_l1(A, B):
    """Relative L1 per component of (nblocks, ncomp, k, j, i) arrays."""
    return [rel_l1(A[:, q], B[:, q]) for q in range(A.shape[1])]


# Reduced versions of the three configurations, sized so the oracle runs
# them in seconds.
def small_configs():
    return {
        "ot64": configs.orszag_tang(64),
        "torus2d_128x64": configs.torus2d(128, 64),
        "torus3d_64x32x32_8blk": configs.torus3d((64, 32, 32), (2, 2, 2)),
    }
