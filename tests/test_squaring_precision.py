"""Why the squarings stay in float64: the s squarings of a CW block's propagator (2 s saturation, s = 11-12
by the kernels' rule) carried out in float32, as a float32 matrix-core product would, move Z by more than
the 1e-5 relative bar of float32 results, while float64 squarings agree with the direct exponential to
1e-12. On the oracle's generator (bmc_matrix) and scipy's expm; CPU only."""
import numpy as np
import scipy.linalg as sl

import bmc_oracle as O
from helpers import four_pool_params


def _z_after_squarings(p, ppm, dtype):
    A, C = O.bmc_matrix(p, 1.0 * 42.577, 0.0, ppm * 42.577 * 3.0)
    X = O.augmented(A, C, 2.0)
    s = int(np.ceil(np.log2(np.abs(X).sum(0).max() / 3.54)))
    E = sl.expm(X / 2.0 ** s).astype(dtype)
    for _ in range(s):
        E = (E @ E).astype(dtype)
    v = np.concatenate([O.initial_state(p), [1.0]])
    zi = 2 * (1 + len(p["cest_k"]))
    z_ref = (sl.expm(X) @ v)[zi]
    return float((E.astype(np.float64) @ v)[zi]), float(z_ref), s


def test_float32_squarings_miss_the_float32_bar_and_float64_ones_do_not():
    p1 = O.make_params(1.3, 0.075, pools=[dict(f=9.01e-4, k=50.0, dw=3.5, t1=1.3, t2=0.1)], b0=3.0)
    worst32 = 0.0
    for p in (p1, four_pool_params()):
        for ppm in (-4.0, 0.5, 3.5):
            z32, z_ref, s = _z_after_squarings(p, ppm, np.float32)
            z64, _, _ = _z_after_squarings(p, ppm, np.float64)
            assert s >= 11
            assert abs(z64 - z_ref) < 1e-12 * abs(z_ref)
            worst32 = max(worst32, abs(z32 - z_ref) / abs(z_ref))
    assert worst32 > 1e-5
