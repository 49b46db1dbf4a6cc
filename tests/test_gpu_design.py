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
