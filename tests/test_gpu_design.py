"""Protocol design through the adjoint kernel: the gradient of the Cramer-Rao cost with respect to a
design that enters the events through torch operations (design/adjoint.py, design/optimize.py).

The reference's tutorial 13 problem: four CW rungs (2 s, 61 offsets) whose amplitudes are the design,
cost averaged over a slow and a fast amide pool. The adjoint gradient must match central differences
of the same cost, and optimize_design must take the same steps with either gradient.
"""
import numpy as np
import pytest

import bmc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    from bmc_jax import _native, settings
    from bmc_jax.fitting import Experiment, FitSpec, FreeParam
    from bmc_jax.simulation import SimParams
    from bmc_jax.simulation.events import SimEvent
    _native.torch_cuda()
    p = O.make_params(1.3, 0.075, pools=[dict(f=9.01e-4, k=50.0, dw=3.5, t1=1.3, t2=0.1)], b0=3.0)
    sp = SimParams(*[np.asarray(p[f], np.float64) if f != "mt_lineshape" else np.int32(0) for f in SimParams._fields])
    m0 = O.initial_state(p)
    rung = O.cw_saturation(np.linspace(-6, 6, 61), b1_uT=1.0, tsat_s=2.0, b0_T=3.0)
    rung = SimEvent(*[np.asarray(rung[k], np.float64) if k != "type" else np.asarray(rung[k], np.int32)
                      for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq", "adc_idx")])

    def build(b1):
        amp = torch.cat([torch.as_tensor(rung.rf_amp) * b1[j] for j in range(4)], dim=0)
        cat = lambda a: np.concatenate([a] * 4, axis=0)
        ev = SimEvent(cat(rung.type), cat(rung.dt), amp, cat(rung.rf_phase), cat(rung.rf_freq), cat(rung.adc_idx))
        return Experiment(sp, ev, m0, 4)

    def build_np(b1):
        cat = lambda a: np.concatenate([a] * 4, axis=0)
        amp = np.concatenate([rung.rf_amp * b1[j] for j in range(4)], axis=0)
        ev = SimEvent(cat(rung.type), cat(rung.dt), amp, cat(rung.rf_phase), cat(rung.rf_freq), cat(rung.adc_idx))
        return Experiment(sp, ev, m0, 4)

    spec = FitSpec([FreeParam("k", "cest_k", 0, 100.0, 5.0, 20000.0, 150.0),
                    FreeParam("f", "cest_f", 0, 0.001, 1e-6, 0.05, 9.01e-4)])
    truths = np.array([[50.0, 9.01e-4], [500.0, 9.01e-4]])
    settings.use_x64(True)
    yield dict(build=build, build_np=build_np, spec=spec, truths=truths, torch=torch)
    settings.use_x64(False)


def test_adjoint_gradient_matches_differences_of_the_cost(setup):
    torch = setup["torch"]
    from bmc_jax.design.adjoint import design_cost
    spec, truths = setup["spec"], setup["truths"]
    d = torch.tensor([0.5, 1.0, 1.5, 2.0], dtype=torch.float64, requires_grad=True)
    c = design_cost(setup["build"](d), spec, 0.002, truths, np.array([0, 1]))
    c.backward()
    g = d.grad.numpy()

    def cost(dv):
        with torch.no_grad():
            return float(design_cost(setup["build"](torch.as_tensor(dv)), spec, 0.002, truths, np.array([0, 1])))
    d0 = np.array([0.5, 1.0, 1.5, 2.0])
    for j in range(4):
        h = 1e-4
        e = np.zeros(4)
        e[j] = h
        fd1 = (cost(d0 + e) - cost(d0 - e)) / (2 * h)
        fd2 = (cost(d0 + e / 2) - cost(d0 - e / 2)) / h
        fd = (4 * fd2 - fd1) / 3
        assert abs(g[j] - fd) <= 1e-6 * np.abs(g).max(), (j, g[j], fd)


def test_cost_matches_the_host_fisher_path(setup):
    """design_cost (batched truths on the GPU, torch inverse) equals the reference formula evaluated with
    fisher_information per truth (optimize.py:80-87)."""
    torch = setup["torch"]
    from bmc_jax.design import fisher_information
    from bmc_jax.design.adjoint import design_cost
    spec, truths = setup["spec"], setup["truths"]
    d0 = np.array([0.5, 1.0, 1.5, 2.0])
    with torch.no_grad():
        c = float(design_cost(setup["build"](torch.as_tensor(d0)), spec, 0.002, truths, np.array([0, 1])))
    exp = setup["build_np"](d0)
    ref = np.mean([np.sum(np.diag(np.linalg.inv(fisher_information(exp, spec, x, 0.002)[0])) / x ** 2)
                   for x in truths])
    assert abs(c - ref) <= 1e-10 * abs(ref)


def test_optimize_design_same_steps_with_either_gradient(setup):
    from bmc_jax.design import optimize_design
    spec, truths = setup["spec"], setup["truths"]
    start, lo, hi = np.array([0.5, 1.0, 1.5, 2.0]), np.full(4, 0.1), np.full(4, 3.0)
    ra = optimize_design(setup["build"], spec, start, lo, hi, 0.002, truths, n_steps=4, lr=0.05, gradient="auto")
    rf = optimize_design(setup["build_np"], spec, start, lo, hi, 0.002, truths, n_steps=4, lr=0.05, gradient="fd",
                         fd_step=1e-5)
    np.testing.assert_allclose(ra.design, rf.design, rtol=1e-6)
    np.testing.assert_allclose(ra.cost, rf.cost, rtol=1e-8)
    assert ra.cost[-1] < ra.cost[0]


# ---- the bound's laws (the reference pins them in tests/test_design.py) --------------------------------

@pytest.fixture(scope="module")
def crlb_setup():
    from bmc_jax import _native
    _native.torch_cuda()
    from bmc_jax.fitting import Experiment, FitSpec, FreeParam
    from bmc_jax.io import example_config
    from bmc_jax.sequences import cw_saturation
    from bmc_jax.simulation import initial_state, to_sim_params
    from bmc_jax.simulation.events import concat_events
    p = example_config("amide")
    jp = to_sim_params(p)
    rung = cw_saturation(np.linspace(-6.0, 6.0, 31), b1_uT=1.0, tsat_s=2.0, b0_T=p.system.b0)

    def ladder(b1):
        ev = concat_events([rung._replace(rf_amp=np.asarray(rung.rf_amp) * b) for b in b1])
        return Experiment(jp, ev, initial_state(jp), p.mz_loc)
    K = FreeParam("k", "cest_k", 0, scale=100.0, lo=5.0, hi=5000.0, x0=150.0)
    F = FreeParam("f", "cest_f", 0, scale=0.001, lo=1e-6, hi=0.05, x0=0.0009)
    return dict(ladder=ladder, K=K, F=F, FitSpec=FitSpec, FreeParam=FreeParam, Experiment=Experiment)


def test_the_bound_scales_with_the_noise(crlb_setup):
    from bmc_jax.design import crlb
    c = crlb_setup
    exp = c["ladder"]((0.5, 1.0, 2.0))
    a, b = crlb(exp, c["FitSpec"]([c["K"], c["F"]]), 0.002), crlb(exp, c["FitSpec"]([c["K"], c["F"]]), 0.004)
    np.testing.assert_allclose(b.sd, 2 * a.sd, rtol=1e-5)
    np.testing.assert_allclose(b.corr, a.corr, atol=1e-6)
    assert a["k"] == a.sd[0] and a.rel_sd[0] == pytest.approx(a.sd[0] / 150.0)


def test_twice_the_data_lowers_the_bound_by_sqrt2(crlb_setup):
    from bmc_jax.design import crlb
    c = crlb_setup
    once = c["ladder"]((0.5, 1.0, 2.0))
    twice = c["Experiment"].stack([once, once])
    a, b = crlb(once, c["FitSpec"]([c["K"], c["F"]]), 0.002), crlb(twice, c["FitSpec"]([c["K"], c["F"]]), 0.002)
    np.testing.assert_allclose(b.sd, a.sd / np.sqrt(2.0), rtol=1e-4)


def test_a_parameter_alone_is_bounded_by_its_own_fisher_entry(crlb_setup):
    from bmc_jax.design import crlb
    c = crlb_setup
    exp = c["ladder"]((0.5, 1.0, 2.0))
    # at the same point (f as the params carry it) the k column is the same column: equal to rounding
    f_cfg = float(np.asarray(exp.params.cest_f)[0])
    both = crlb(exp, c["FitSpec"]([c["K"], c["F"]]), 0.002, x=[150.0, f_cfg])
    alone = crlb(exp, c["FitSpec"]([c["K"]]), 0.002)
    assert alone.sd[0] == pytest.approx(1.0 / np.sqrt(both.fisher[0, 0]), rel=1e-6)
    assert alone.sd[0] < both.sd[0]


def test_the_bound_needs_one_sample_and_plain_values(crlb_setup):
    from bmc_jax.design import crlb
    from bmc_jax.simulation import tile_params
    c = crlb_setup
    exp = c["ladder"]((0.5, 1.0, 2.0))
    with pytest.raises(ValueError, match="one sample"):
        crlb(c["Experiment"](tile_params(exp.params, 2), exp.events, exp.m0, exp.mz_loc),
             c["FitSpec"]([c["K"], c["F"]]), 0.002)
    shared = c["FreeParam"]("r2w", "water_r2", None, scale=1.0, lo=0.5, hi=2.0, x0=1.0, scope="shared", mode="factor")
    with pytest.raises(ValueError, match="plain values"):
        crlb(exp, c["FitSpec"]([c["K"], c["F"], shared]), 0.002)
