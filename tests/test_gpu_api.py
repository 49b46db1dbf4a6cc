"""The drop-in Python API on the GPU, tested the way the reference tests itself.

Physics invariants (reference tests/test_solver.py), padding and stacking to the last bit
(test_experiment.py), recovery of a self-simulated truth (test_fitting.py), and the LM loop against
the float64 oracle's own LM, accept for accept.
"""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC

pytestmark = pytest.mark.gpu

B0_T = 2.8936
PPM = np.linspace(-6.0, 6.0, 121)
I_POS = int(np.argmin(np.abs(PPM - 3.5)))
I_NEG = PPM.size - 1 - I_POS


@pytest.fixture(scope="module")
def bj():
    import bmc_jax
    from bmc_jax import _native
    _native.torch_cuda()
    bmc_jax.settings.use_x64(False)
    return bmc_jax


def params(bj, f=0.005, k=300.0, dw=3.5, t2_pool_s=0.01, b0=B0_T):
    """tests/test_solver.py:30-39 of the reference: one water and one CEST pool, float32."""
    from bmc_jax.simulation import SimParams
    s = np.float32
    return SimParams(water_r1=s(1 / 3.0), water_r2=s(1 / 0.6), water_f=s(1.0), cest_k=np.float32([k]),
                     cest_f=np.float32([f]), cest_dw=np.float32([dw]), cest_r1=np.float32([1.0]),
                     cest_r2=np.float32([1.0 / t2_pool_s]), mt_k=s(0), mt_f=s(0), mt_dw=s(0), mt_r1=s(0), mt_r2=s(1),
                     mt_lineshape=np.int32(0), b0=s(b0), gamma=s(267.5153), b0_inhom=s(0), rel_b1=s(1.0))


def z_spectrum(bj, p, ppm=PPM, b1_uT=1.0, tsat_s=2.0, b0=B0_T):
    from bmc_jax.sequences import cw_saturation
    from bmc_jax.simulation import initial_state, simulate_spectrum
    ev = cw_saturation(ppm, b1_uT=b1_uT, tsat_s=tsat_s, b0_T=b0)
    return np.asarray(simulate_spectrum(p, ev, initial_state(p), return_full_trajectory=False)[:, 4])


def test_solver_invariants(bj):
    z = z_spectrum(bj, params(bj))
    assert 0.95 < z[0] < 1.01
    assert abs(PPM[int(np.argmin(z))]) <= 0.5 and z.min() < z[I_POS]
    sym = z_spectrum(bj, params(bj, f=0.0))
    assert np.abs(sym - sym[::-1]).max() < 2e-3
    assert z[I_NEG] - z[I_POS] > 0.05
    p = params(bj)
    assert z_spectrum(bj, p, b1_uT=2.0).mean() < z.mean()
    assert z_spectrum(bj, p, tsat_s=4.0).mean() < z_spectrum(bj, p, tsat_s=1.0).mean()
    k0 = z_spectrum(bj, params(bj, k=0.0))
    assert abs(k0[I_NEG] - k0[I_POS]) < 2e-3


@pytest.mark.parametrize("b0", [1.5, 2.8936, 7.0])
def test_ppm_axis_scales_with_field(bj, b0):
    ppm = np.linspace(0.5, 6.0, 45)
    zw = z_spectrum(bj, params(bj, b0=b0), ppm=ppm, b0=b0)
    zn = z_spectrum(bj, params(bj, f=0.0, b0=b0), ppm=ppm, b0=b0)
    assert abs(ppm[int(np.argmax(zn - zw))] - 3.5) < 0.4


def _experiment_at(bj, tsat_s, b1_uT=1.0, name="amide"):
    from bmc_jax.fitting import Experiment
    from bmc_jax.io import example_config
    from bmc_jax.sequences import cw_saturation
    from bmc_jax.simulation import initial_state, to_sim_params
    p = example_config(name)
    jp = to_sim_params(p)
    ev = cw_saturation(np.linspace(-5.0, 5.0, 21), b1_uT=b1_uT, tsat_s=tsat_s, b0_T=p.system.b0)
    return Experiment(jp, ev, initial_state(jp), p.mz_loc)


def test_padding_and_stacking_are_exact(bj):
    """test_experiment.py:29-102: front padding, equal and ragged stacking change nothing, to the last bit."""
    from bmc_jax.fitting import Experiment
    from bmc_jax.simulation import pad_events
    e = _experiment_at(bj, 2.0)
    z = np.asarray(e.simulate())
    padded = Experiment(e.params, pad_events(e.events, e.n_events + 37), e.m0, e.mz_loc)
    assert np.abs(z - np.asarray(padded.simulate())).max() == 0.0
    parts = [_experiment_at(bj, 2.0, b1_uT=b) for b in (0.5, 1.0, 2.0)]
    joint = Experiment.stack(parts)
    zj = np.asarray(joint.simulate())
    for i, part in enumerate(parts):
        assert np.abs(zj[i * 21:(i + 1) * 21] - np.asarray(part.simulate())).max() == 0.0
    ragged = [Experiment(p.params, pad_events(p.events, p.n_events + 2 * i), p.m0, p.mz_loc)
              for i, p in enumerate([_experiment_at(bj, t) for t in (0.5, 2.0, 4.0)])]
    zr = np.asarray(Experiment.stack(ragged).simulate())
    for i, part in enumerate(ragged):
        assert np.abs(zr[i * 21:(i + 1) * 21] - np.asarray(part.simulate())).max() == 0.0
    with pytest.raises(ValueError, match="mz_loc"):
        Experiment.stack([_experiment_at(bj, 2.0), _experiment_at(bj, 2.0, name="amide_noe")])
    with pytest.raises(ValueError, match="cannot pad"):
        pad_events(e.events, e.n_events - 1)


def test_trajectory_ends_at_the_final_state(bj):
    from bmc_jax.simulation import simulate_spectrum
    e = _experiment_at(bj, 2.0)
    traj = simulate_spectrum(e.params, e.events, e.m0)          # default: the full trajectory
    fin = simulate_spectrum(e.params, e.events, e.m0, return_full_trajectory=False)
    assert traj.shape == (21, 2, 7)
    np.testing.assert_allclose(traj[:, -1], fin, rtol=0, atol=1e-7)


LADDER = (0.5, 1.0, 1.5, 2.0)
K_TRUE = np.linspace(30.0, 400.0, 8)


@pytest.fixture(scope="module")
def fitted(bj):
    """test_fitting.py:29-60: eight samples differing in k, a 4-rung QUESP ladder, fit_batch n_iter 30."""
    from bmc_jax.fitting import FitSpec, FreeParam, fit_batch
    from bmc_jax.io import example_config
    from bmc_jax.sequences import cw_saturation
    from bmc_jax.simulation import SimParams, initial_state, simulate_voxels, to_sim_params
    from bmc_jax.simulation.events import concat_events
    p = example_config("amide")
    jp = to_sim_params(p)
    ppm = np.linspace(-6.0, 6.0, 41)
    events = concat_events([cw_saturation(ppm, b1_uT=b, tsat_s=2.0, b0_T=p.system.b0) for b in LADDER])
    n = K_TRUE.size
    base = SimParams(*[np.broadcast_to(np.asarray(a), (n,) + np.shape(a)) for a in jp])
    base = base._replace(cest_k=np.asarray(K_TRUE[:, None], np.float32))
    M0 = initial_state(base)
    Z = simulate_voxels(base, events, M0)[..., p.mz_loc]
    spec = FitSpec([FreeParam("k", "cest_k", 0, scale=100.0, lo=5.0, hi=2000.0, x0=150.0),
                    FreeParam("f", "cest_f", 0, scale=0.001, lo=1e-5, hi=0.01, x0=0.0005)])
    X = np.asarray(fit_batch(base, events, M0, p.mz_loc, spec, Z, n_iter=30, skip_first=False, base_batched=True))
    return p, X, spec


def test_exchange_rate_and_fraction_are_recovered(fitted):
    p, X, spec = fitted
    rel = np.abs(X[:, 0] / K_TRUE - 1.0)
    assert rel.max() < 0.05
    assert rel[K_TRUE > 100].max() < 0.005
    assert np.abs(X[:, 1] / p.cest_pools[0].f - 1.0).max() < 0.03
    flux = np.abs((X[:, 0] * X[:, 1]) / (K_TRUE * p.cest_pools[0].f) - 1.0).max()
    assert flux <= rel.max() + 1e-9
    assert (X[:, 0] >= spec.lo_phys[0]).all() and (X[:, 0] <= spec.hi_phys[0]).all()


def test_readme_quick_start(bj):
    """The reference's README quick start (README.md:61-94), run as printed: an amide QUESP ladder simulated
    and fitted back with fit(). The README prints k = 49.617, f = 0.00090772, rms 2.06e-05 against the truth
    k = 50, f = 9.01e-4, and says the last digits depend on device and precision; the band holds both."""
    from bmc_jax.fitting import Experiment, FitSpec, FreeParam, fit
    from bmc_jax.io import example_config
    from bmc_jax.sequences import cw_saturation
    from bmc_jax.simulation import initial_state, to_sim_params
    params = example_config("amide")
    jp = to_sim_params(params)
    m0 = initial_state(jp)
    ppm = np.linspace(-6, 6, 61)
    ladder = Experiment.stack([
        Experiment(jp, cw_saturation(ppm, b1_uT=b, tsat_s=2.0, b0_T=params.system.b0), m0, params.mz_loc)
        for b in (0.5, 1.0, 2.0, 4.0)])
    z = np.asarray(ladder.simulate())
    spec = FitSpec([FreeParam("k", "cest_k", 0, scale=100.0, lo=5.0, hi=5000.0, x0=300.0),
                    FreeParam("f", "cest_f", 0, scale=0.001, lo=1e-6, hi=0.05, x0=0.0005)])
    res = fit(ladder, spec, z[None, :], n_per_block=len(ppm))
    k, f = float(res["k"][0]), float(res["f"][0])
    assert 49.0 <= k <= 50.5, res.summary()
    assert 8.9e-4 <= f <= 9.15e-4, res.summary()
    assert float(res.rms[0]) <= 5e-5, res.summary()
    assert "sample" in res.summary() and "rms" in res.summary()


def test_levenberg_marquardt_without_physics(bj):
    from bmc_jax.fitting import levenberg_marquardt
    r = lambda x: np.array([10.0 * (x[1] - x[0] ** 2), 1.0 - x[0]])
    x = levenberg_marquardt(r, np.array([-1.2, 1.0]), np.array([-5.0, -5.0]), np.array([5.0, 5.0]), n_iter=200)
    assert np.allclose(x, [1.0, 1.0], atol=1e-3)
    r2 = lambda x: np.array([x[0] - 1.0, x[1] - 2.0, 0.1 * x[0] * x[1]])
    held = levenberg_marquardt(r2, np.array([0.0, 5.0]), np.array([-10.0, -10.0]), np.array([10.0, 10.0]),
                               n_iter=30, free=np.array([1.0, 0.0]))
    assert float(held[1]) == 5.0


def test_gpu_lm_follows_the_oracle_lm_step_for_step(bj):
    """float64: the GPU-resident LM and the oracle's LM (lm.py restated, C-oracle residual and Jacobian)
    take the same accept/reject decisions and land on the same point."""
    from bmc_jax import settings
    from bmc_jax.fitting import FitSpec, FreeParam, fit_batch
    from bmc_jax.simulation import SimParams
    settings.use_x64(True)
    try:
        p = O.make_params(1.3, 0.075, pools=[dict(f=9.01e-4, k=50.0, dw=3.5, t1=1.3, t2=0.1)], b0=3.0)
        ppm = np.linspace(-6, 6, 31)
        ev = {k: np.concatenate([v for _ in range(3)]) for k, v in O.cw_saturation(ppm, 1.0, 2.0, 3.0).items()}
        ev["rf_amp"] = ev["rf_amp"] * np.repeat([0.5, 1.0, 2.0], 31)[:, None].astype(np.float32)
        m0 = O.initial_state(p)
        idx = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1)]
        z_true = OC.simulate(OC.pack(p)[None], 1, ev, m0)[0][0, :, 4]
        scales = np.array([100.0, 0.001])

        def resid(xs):
            q = OC.pack(p).copy(); q[idx] = xs * scales
            return OC.simulate(q[None], 1, ev, m0)[0][0, :, 4] - z_true

        def jac(xs):
            q = OC.pack(p).copy(); q[idx] = xs * scales
            return OC.simulate(q[None], 1, ev, m0, slots=idx)[1][0, :, :, 4] * scales

        x0 = np.array([2.0, 0.4])
        xo, Xo, Co = O.levenberg_marquardt(resid, jac, x0, [0.05, 0.01], [20.0, 10.0], n_iter=15, trajectory=True)
        sp = SimParams(*[np.asarray(p[f], np.float64) if f != "mt_lineshape" else np.int32(p[f])
                         for f in SimParams._fields])
        spec = FitSpec([FreeParam("k", "cest_k", 0, 100.0, 5.0, 2000.0, 200.0),
                        FreeParam("f", "cest_f", 0, 0.001, 1e-5, 0.01, 0.0004)])
        from bmc_jax.simulation.events import SimEvent
        evs = SimEvent(ev["type"], ev["dt"].astype(np.float64), ev["rf_amp"].astype(np.float64),
                       ev["rf_phase"].astype(np.float64), ev["rf_freq"].astype(np.float64), ev["adc_idx"])
        X = fit_batch(sp, evs, m0, 4, spec, z_true[None], n_iter=15, skip_first=False)
        np.testing.assert_allclose(X[0] / scales, xo, rtol=1e-7)
    finally:
        settings.use_x64(False)


def test_fit_returns_a_fitresult(bj):
    from bmc_jax.fitting import FitSpec, FreeParam, fit
    e = _experiment_at(bj, 2.0)
    z = np.asarray(e.simulate())
    spec = FitSpec([FreeParam("k", "cest_k", 0, 100.0, 5.0, 2000.0, 150.0),
                    FreeParam("f", "cest_f", 0, 0.001, 1e-5, 0.01, 0.0005)])
    res = fit(e, spec, z, n_iter=10)
    assert res.x.shape == (1, 2) and res.z_fit.shape == (1, 21)
    assert np.isfinite(res.rms).all()
    assert res["k"].shape == (1,)
    with pytest.raises(ValueError, match="fit_joint"):
        fit(e, FitSpec([FreeParam("r2w", "water_r2", None, 10.0, 1.0, 100.0, 14.0, scope="shared")]), z)
