"""The reference's tutorials 08 and 09 on its shipped phantom measurement (tests/golden/phantom, written by
tests/golden/make_phantom.py): six tubes, the converted scanner sequences (4041-5051 events per row), the
drop-in fit() with 60 LM iterations.

Checked: the model on these sequences against the float64 C oracle (rows of the real ladder, fitted
values of a tube), and the fits themselves against what the phantom's recipe fixes regardless of any
calibration: the two GAG tubes (100 and 50 mg/mL of one stock) must give a pool-size ratio of 2."""
import numpy as np
import pytest

import oracle_c as OC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def quesp():
    from bmc_jax import _native
    _native.torch_cuda()
    import bench
    return bench.phantom_experiment("quesp")


def test_the_ladder_model_matches_the_oracle(quesp):
    """Z of two tubes on rows of three rungs (pulses at accumulated phases: kept and reused by the plan)
    against the C oracle, event by event, float64."""
    from bmc_jax import _engine
    from bmc_jax.simulation.events import SimEvent, slice_rows
    from bmc_jax.simulation.params import SimParams, pack
    ladder, spec, Z, noise, names, npb = quesp
    rows = np.array([3, npb + 40, 5 * npb + 77])
    ev = slice_rows(ladder.events, rows)
    ev64 = SimEvent(*[np.asarray(a, np.float64) if np.asarray(a).dtype.kind == "f" else np.asarray(a) for a in ev])
    for t in (0, 3):
        p = SimParams(*[np.asarray(a)[t:t + 1] for a in ladder.params])
        p64 = SimParams(*[np.asarray(a, np.float64) if np.asarray(a).dtype.kind == "f" else np.asarray(a) for a in p])
        m0 = np.asarray(ladder.m0[t:t + 1], np.float64)
        z = _engine.simulate_states(p64, ev64, m0, n_voxels=1, dtype=np.float64)[0, :, 4]
        evd = {k: np.asarray(getattr(ev64, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
        zo = OC.simulate(pack(p64, 1, dtype=np.float64), 1, evd, m0)[0][0, :, 4]
        np.testing.assert_allclose(z, zo, rtol=0, atol=1e-11)


def test_tutorial_08_fits_the_phantom(quesp):
    from bmc_jax.fitting import fit
    import bench
    ladder, spec, Z, noise, names, npb = quesp
    res = fit(ladder, spec, Z, n_iter=60, n_per_block=npb)
    assert np.all(np.isfinite(res.x)) and np.all(res.x >= spec.lo_phys) and np.all(res.x <= spec.hi_phys)
    assert np.all(res.rms / noise < 15.0), res.rms / noise
    pairs = {p["pair"]: p["f_ratio"] for p in bench.phantom_pairs(np.asarray(res["f"]), names)}
    assert pairs["GAG 100 / 50 mg/mL"] == pytest.approx(2.0, rel=0.1), pairs
    # the fast-exchanging GAG pools are measured, not railed
    gag = [names.index("GAG100_pH6.31"), names.index("GAG50_pH6.20")]
    assert not res.at_bounds[gag].any()


def test_tutorial_09_fits_the_phantom():
    from bmc_jax.fitting import fit
    import bench
    ladder, spec, Z, noise, names, npb = bench.phantom_experiment("quest")
    res = fit(ladder, spec, Z, n_iter=60, n_per_block=npb)
    assert np.all(np.isfinite(res.x))
    assert np.all(res.rms / noise < 15.0), res.rms / noise
    pairs = {p["pair"]: p["f_ratio"] for p in bench.phantom_pairs(np.asarray(res["f"]), names)}
    assert pairs["GAG 100 / 50 mg/mL"] == pytest.approx(2.0, rel=0.1), pairs


def test_tutorial_10_shares_the_water_over_seven_tubes():
    """fit_joint over the seven tubes (PBS included) with the water R2 factor, M_start and a 15-node B0
    spread shared: the joint cost falls below the start's, the shared values stay inside their bounds,
    and the GAG pair keeps its ratio of 2."""
    import bench
    from dataclasses import replace
    from bmc_jax.fitting import FitSpec, FreeParam, fit, fit_joint
    from bmc_jax.simulation import B0Spread
    ladder6, spec3, Z6, _, names6, npb = bench.phantom_experiment("quesp", "pool")
    ladder, _, Z, noise, names, _ = bench.phantom_experiment("quesp", "all")
    pool = [names.index(nm) for nm in names6]
    plain = fit(ladder6, spec3, Z6, n_iter=60, n_per_block=npb)
    Kp, Fp, T2p = (replace(sl, samples=tuple(pool)) for sl in spec3.slots)
    spec = FitSpec([Kp, Fp, T2p,
                    FreeParam("r2w", "water_r2", None, 1.0, 0.5, 2.0, 1.0, scope="shared", mode="factor"),
                    FreeParam("m_start", "m_start", None, 1.0, 0.2, 1.0, 0.557, scope="shared"),
                    FreeParam("sigma", "b0_spread", None, 10.0, 0.0, 8.0, 5.0, scope="shared")])
    x0 = np.tile([sl.x0 for sl in spec.slots], (len(names), 1))
    x0[pool, :3] = plain.x
    res = fit_joint(ladder, spec, Z, n_iter=12, n_per_block=npb, x0_phys=x0, spread=B0Spread(nodes=15))
    assert res.cost[-1] < res.cost[0] and res.nan_steps == 0
    for name in ("r2w", "m_start", "sigma"):
        v = res[name]
        assert np.all(v == v[0]) and np.isfinite(v[0])
    pairs = {p["pair"]: p["f_ratio"] for p in bench.phantom_pairs(np.asarray(res["f"]), names)}
    assert pairs["GAG 100 / 50 mg/mL"] == pytest.approx(2.0, rel=0.1), pairs
    assert not res.free[names.index("PBS"), 0]          # PBS has no pool: its k is held
