"""MRF dictionaries (src/bmc_jax/mrf/dictionary.py): the host grid expansion (CPU), and the fingerprints
of a small grid against the float64 C oracle, chunked and unchunked (GPU)."""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import cfg1_events


def base_params(nc=1):
    from bmc_jax.simulation import SimParams
    a = lambda v: np.asarray(v, np.float64)
    return SimParams(water_r1=a(1 / 1.3), water_r2=a(1 / 0.075), water_f=a(1.0), cest_k=a([50.0] * nc),
                     cest_f=a([9.01e-4] * nc), cest_dw=a([3.5] * nc), cest_r1=a([1 / 1.3] * nc),
                     cest_r2=a([10.0] * nc), mt_k=a(0.0), mt_f=a(0.0), mt_dw=a(0.0), mt_r1=a(0.0), mt_r2=a(1.0),
                     mt_lineshape=np.int32(0), b0=a(3.0), gamma=a(267.5153), b0_inhom=a(0.0), rel_b1=a(1.0))


def test_grid_is_the_cartesian_product_first_field_slowest():
    from bmc_jax.mrf import build_param_grid
    g = build_param_grid(base_params(), {"water_r1": [0.5, 1.0], "cest_k": [10.0, 20.0, 30.0]})
    np.testing.assert_array_equal(g.water_r1, [0.5, 0.5, 0.5, 1.0, 1.0, 1.0])
    np.testing.assert_array_equal(g.cest_k[:, 0], [10.0, 20.0, 30.0] * 2)
    assert g.cest_f.shape == (6, 1) and g.b0.shape == (6,) and np.all(g.cest_f == 9.01e-4)


def test_grid_refuses_what_it_cannot_sweep():
    from bmc_jax.mrf import build_param_grid
    with pytest.raises(ValueError, match="at least one"):
        build_param_grid(base_params(), {})
    with pytest.raises(ValueError, match="more than one element"):
        build_param_grid(base_params(nc=2), {"cest_k": [1.0, 2.0]})
    with pytest.raises(ValueError, match="not SimParams fields"):
        build_param_grid(base_params(), {"k": [1.0]})


def test_the_package_exports():
    import bmc_jax.mrf as M
    assert M.__all__ == ["build_param_grid", "generate_dictionary"]
    assert all(hasattr(M, n) for n in M.__all__)


@pytest.mark.gpu
def test_fingerprints_match_the_oracle():
    from bmc_jax.mrf import generate_dictionary
    from bmc_jax.simulation import SimEvent, initial_state
    from bmc_jax import settings
    ev = cfg1_events()
    seq = SimEvent(*[np.asarray(ev[k]) if k in ev else np.where(ev["type"] == 3, 0, -1).astype(np.int32)
                     for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq", "adc_idx")])
    base = base_params()
    grid = {"cest_k": [20.0, 200.0, 2000.0], "water_r1": [0.5, 0.8], "cest_f": [1e-3, 5e-3]}
    m0 = initial_state(base)
    settings.use_x64(True)
    try:
        entries, fp = generate_dictionary(base, seq, m0, grid, water_z_index=4)
        _, fp_chunked = generate_dictionary(base, seq, m0, grid, water_z_index=4, chunk_size=5)
    finally:
        settings.use_x64(False)
    assert fp.shape == (12, 61) and fp.dtype == np.float32
    np.testing.assert_array_equal(fp, fp_chunked)
    for i in range(12):
        p = O.make_params(1 / entries["water_r1"][i], 0.075,
                          pools=[dict(f=entries["cest_f"][i], k=entries["cest_k"][i], dw=3.5, t1=1.3, t2=0.1)], b0=3.0)
        m, _ = OC.simulate(OC.pack(p)[None], 1, ev, np.asarray(m0))
        np.testing.assert_allclose(fp[i], m[0, :, 4], rtol=0, atol=1e-6)   # float32 fingerprints
