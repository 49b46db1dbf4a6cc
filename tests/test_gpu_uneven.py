"""Uneven batched LM (bmc_ubatch_*): problems with different pool counts, rows and free values in one
persistent launch per sweep.

Parity: every problem's Z, Jacobian and LM result must equal what the uniform path (bmc_simulate /
bmc_lm_fit on that problem alone, itself pinned to the float64 oracle) gives, whatever launch mapping
the autotuner picks; noise-free problems must be recovered.
"""
import numpy as np
import pytest

import bmc_oracle as O

pytestmark = pytest.mark.gpu


def _sim_params(p):
    from bmc_jax.simulation import SimParams
    return SimParams(*[np.asarray(p[f], np.float64) if f != "mt_lineshape" else np.int32(p[f])
                       for f in SimParams._fields])


def make_problem(rng, nc, mt, n_rungs, n_off, P):
    from bmc_jax.fitting import FitSpec, FreeParam
    from bmc_jax.fitting.uneven import UnevenProblem
    from bmc_jax.simulation.events import SimEvent
    pools = [dict(f=float(rng.uniform(1e-3, 1e-2)), k=float(np.exp(rng.uniform(np.log(20), np.log(2000)))),
                  dw=float(rng.uniform(1.0, 4.0)), t1=1.3, t2=0.05) for _ in range(nc)]
    mtd = dict(f=0.05, k=20.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape="lorentzian") if mt else None
    p = O.make_params(float(rng.uniform(0.9, 2.0)), float(rng.uniform(0.04, 0.12)), pools=pools, mt=mtd, b0=3.0,
                      b0_inhom=float(rng.uniform(-0.1, 0.1)))
    b1s = rng.choice([0.5, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 8.0], size=n_rungs, replace=False)
    evs = [O.cw_saturation(np.linspace(-5, 5, n_off), b1_uT=float(b), tsat_s=1.0, b0_T=3.0) for b in b1s]
    ev = {k: np.concatenate([e[k] for e in evs], axis=0) for k in evs[0]}
    events = SimEvent(*[np.asarray(ev[k]) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq", "adc_idx")])
    cands = []
    for i in range(nc):
        cands += [FreeParam(f"k{i}", "cest_k", i, 100.0, 1.0, 1e4, pools[i]["k"]),
                  FreeParam(f"f{i}", "cest_f", i, 0.001, 1e-5, 0.05, pools[i]["f"])]
    cands += [FreeParam("r1w", "water_r1", None, 1.0, 0.1, 3.0, float(p["water_r1"])),
              FreeParam("r2w", "water_r2", None, 10.0, 1.0, 60.0, float(p["water_r2"])),
              FreeParam("db0", "b0_inhom", None, 0.1, -0.5, 0.5, float(p["b0_inhom"]))]
    if mt:
        cands += [FreeParam("kmt", "mt_k", None, 10.0, 1.0, 100.0, 20.0), FreeParam("fmt", "mt_f", None, 0.01, 1e-3, 0.3, 0.05)]
    pick = sorted(rng.choice(len(cands), size=min(P, len(cands)), replace=False))
    spec = FitSpec([cands[j] for j in pick])
    sp = _sim_params(p)
    return UnevenProblem(params=sp, events=events, m0=O.initial_state(p), spec=spec, z_meas=np.zeros(events.dt.shape[0]),
                         x0_phys=spec.x0_phys * rng.uniform(0.8, 1.25, spec.n))


@pytest.fixture(scope="module")
def problems():
    from bmc_jax import _native
    _native.torch_cuda()
    rng = np.random.default_rng(12)
    shapes = [(1, False, 1, 9, 2), (2, True, 2, 7, 5), (3, False, 3, 5, 4), (4, True, 1, 11, 8), (1, True, 4, 6, 3),
              (2, False, 2, 13, 6), (4, False, 2, 5, 2), (3, True, 1, 9, 7), (1, False, 2, 8, 4), (2, True, 3, 6, 3),
              (4, True, 3, 4, 5), (1, True, 1, 31, 2), (3, False, 2, 10, 6)]
    probs = [make_problem(rng, *s) for s in shapes]
    # measured Z: each problem simulated at its own true values (noise-free)
    from bmc_jax.simulation.simulate import simulate_spectrum
    for p in probs:
        p.z_meas = np.asarray(simulate_spectrum(p.params, p.events, p.m0, return_full_trajectory=False))[:, 2 * (1 + len(p.params.cest_k))]
    return probs


def _uniform_sim(p, x_scaled):
    """The uniform path on one problem with its forward-mode Jacobian kernel, the arithmetic the uneven
    sweep shares when asked for forward mode (single-event rows otherwise take the reverse kernel, equal
    to rounding: tests/test_cw_reverse.py)."""
    from bmc_jax import _engine
    from bmc_jax.fitting.params import write_params
    from bmc_jax.fitting.residual import spec_slots
    nc = len(p.params.cest_k)
    pw = write_params(p.params, x_scaled, p.spec)
    m, J = _engine.simulate_with_jac(pw, p.events, p.m0, spec_slots(p.spec, nc), p.spec.scales, 2 * (1 + nc),
                                     n_voxels=1, dtype=np.float64, forward_jacobian=True)
    return m[0, :, 2 * (1 + nc)], J[0]


@pytest.mark.parametrize("cfg", [(1, 1024, 1), (2, 256, 0), (1, 7, 0)])
def test_simulate_matches_the_uniform_kernel(problems, cfg):
    from bmc_jax.fitting.uneven import UnevenBatch
    b = UnevenBatch(problems, dtype=np.float64)
    b.configure(*cfg)
    x = b.x0_scaled()
    z, J = b.simulate(x, forward_jacobian=True)
    z = z.cpu().numpy()
    for i, p in enumerate(problems):
        zu, Ju = _uniform_sim(p, x[b.x_off[i]:b.x_off[i + 1]])
        zi = z[b.row0[i]:b.row0[i] + b.n_rows[i]]
        np.testing.assert_allclose(zi, zu, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(b.jacobian_of(J, i), Ju, rtol=1e-12, atol=1e-14 * np.abs(Ju).max())


def test_lm_matches_the_uniform_lm(problems):
    from bmc_jax.fitting.independent import _lm_batched
    from bmc_jax.fitting.uneven import UnevenBatch
    from bmc_jax.simulation.voxels import tile_params
    b = UnevenBatch(problems, dtype=np.float64)
    b.configure(2, 64, 1)
    x, cost, nd = b.lm_fit(n_iter=6, forward_jacobian=True)
    xs = b.split(x)
    for i, p in enumerate(problems):
        nc = len(p.params.cest_k)
        x0 = p.spec.to_scaled(p.x0_phys)[None]
        xu, _, _, _ = _lm_batched(tile_params(p.params, 1), p.events, p.m0[None], 2 * (1 + nc), p.spec,
                                  np.asarray(p.z_meas)[None], x0, 6, np.float64, forward_jacobian=True)
        np.testing.assert_allclose(xs[i], xu[0] * p.spec.scales, rtol=1e-11)
    assert np.all(nd == 6)


def test_graph_and_early_exit_and_recovery(problems):
    from bmc_jax.fitting.uneven import UnevenBatch, autotune
    b = UnevenBatch(problems, dtype=np.float64)
    best, times = autotune(b, candidates=[(1, 256, 0), (2, 512, 1)], repeats=1)
    assert best in times
    x1, c1, _ = b.lm_fit(n_iter=40, graph=True)
    x2, c2, _ = b.lm_fit(n_iter=40, graph=False)
    np.testing.assert_array_equal(x1, x2)
    x3, c3, nd3 = b.lm_fit(n_iter=40, early_exit=True, early_rms=1e-7, early_ftol=0.0)
    assert np.all(nd3 <= 40) and np.all(nd3 >= 1)
    z0, _ = b.simulate(b.x0_scaled(), jacobian=False)
    r0 = z0.cpu().numpy() - np.concatenate([p.z_meas for p in problems])
    c0 = np.array([np.sum(r0[b.row0[i]:b.row0[i] + b.n_rows[i]] ** 2) for i in range(b.n)])
    assert np.all(c3 <= c0)
    rms = np.sqrt(c3 / b.n_rows)
    assert np.median(rms) <= 1e-5, rms


def test_reverse_jacobian_sweep_matches_forward_mode(problems):
    """The uneven batch's Jacobian sweep of single-event rows by reverse accumulation (one item per row,
    every tangent at once) against its forward-mode sweep (items per tangent pass)."""
    from bmc_jax.fitting.uneven import UnevenBatch
    b = UnevenBatch(problems, dtype=np.float64)
    b.configure(2, 256, 1)
    x = b.x0_scaled()
    z_r, J_r = b.simulate(x)
    z_r = z_r.cpu().numpy()
    z_f, J_f = b.simulate(x, forward_jacobian=True)
    # different Taylor degree / squaring splits of exp(X): equal to rounding, the oracle's 1e-11 bar
    np.testing.assert_allclose(z_r, z_f.cpu().numpy(), rtol=0, atol=1e-11)
    for i in range(len(problems)):
        jr, jf = b.jacobian_of(J_r, i), b.jacobian_of(J_f, i)
        assert np.max(np.abs(jr - jf)) <= 1e-9 * np.abs(jf).max(), i
