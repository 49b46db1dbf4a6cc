"""Host-side logic of the package that needs no GPU: the joint LM step with active bounds, the B0-spread
quadrature, the host Levenberg-Marquardt's held entries, FitSpec validation (the reference pins these in
tests/test_joint.py and tests/test_fitting.py), and the MLP kernels' GELU fit against scipy."""
import numpy as np
import pytest


def test_active_bounds_hold_a_value_that_would_leave():
    from bmc_jax.fitting.joint import _lm_trial
    rng = np.random.default_rng(1)
    J = rng.normal(size=(30, 3))
    r = rng.normal(size=30)
    if (J.T @ r)[0] < 0:                     # the gradient of value 0 pushes it below its bound
        J[:, 0] *= -1
    x = np.array([0.0, 0.5, 0.5])
    lo, hi = np.zeros(3), np.full(3, 10.0)
    lam = 0.1
    held = np.asarray(_lm_trial(x, lam, J, r, lo, hi, active_bounds=True))
    Jr = J[:, 1:]
    H, g = Jr.T @ Jr, Jr.T @ r
    want = x[1:] + np.linalg.solve(H + lam * (np.diag(np.diag(H)) + 1e-9 * np.eye(2)), -g)
    assert held[0] == 0.0
    np.testing.assert_allclose(held[1:], want, rtol=1e-10)
    clipped = np.asarray(_lm_trial(x, lam, J, r, lo, hi))
    assert clipped[0] == 0.0 and not np.allclose(clipped[1:], want, rtol=1e-4)


def test_spread_nodes_and_weights():
    from bmc_jax.simulation import B0Spread
    u, w = B0Spread(shape="gauss", nodes=5).nodes_weights()
    assert w.sum() == pytest.approx(1.0) and np.sum(w * u ** 2) == pytest.approx(1.0)
    u, w = B0Spread(shape="lorentz", nodes=9, cut=20.0).nodes_weights()
    T = np.arctan(20.0)
    assert np.sum(w / (1 + u ** 2)) == pytest.approx(0.5 + np.sin(2 * T) / (4 * T), rel=1e-10)
    assert np.all(w > 0) and np.allclose(u, -u[::-1])
    with pytest.raises(ValueError, match="shape"):
        B0Spread(shape="box")


def test_levenberg_marquardt_holds_what_is_not_free():
    from bmc_jax.fitting import levenberg_marquardt
    resid = lambda x: np.array([x[0] - 1.0, x[1] - 2.0, 0.1 * x[0] * x[1]])
    x0, lo, hi = np.array([0.0, 5.0]), np.array([-10.0, -10.0]), np.array([10.0, 10.0])
    held = levenberg_marquardt(resid, x0, lo, hi, n_iter=30, free=np.array([1.0, 0.0]))
    assert float(held[1]) == 5.0
    both = levenberg_marquardt(resid, x0, lo, hi, n_iter=30)
    ones = levenberg_marquardt(resid, x0, lo, hi, n_iter=30, free=np.array([1.0, 1.0]))
    np.testing.assert_array_equal(np.asarray(both), np.asarray(ones))


def test_levenberg_marquardt_works_without_any_physics():
    """Rosenbrock as residuals (10 (y - x^2), 1 - x): the minimum (1, 1)."""
    from bmc_jax.fitting import levenberg_marquardt
    resid = lambda x: np.array([10.0 * (x[1] - x[0] ** 2), 1.0 - x[0]])
    x = levenberg_marquardt(resid, np.array([-1.2, 1.0]), np.array([-5.0, -5.0]), np.array([5.0, 5.0]), n_iter=200)
    np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-6)


def test_free_param_validation():
    from bmc_jax.fitting import FreeParam
    with pytest.raises(ValueError, match="b0_spread"):
        FreeParam("s", "b0_spread", None, 1.0, 0.0, 1.0, 0.5, mode="factor")
    with pytest.raises(ValueError, match="scope"):
        FreeParam("k", "cest_k", 0, 1.0, 0.0, 1.0, 0.5, scope="voxel")
    with pytest.raises(ValueError, match="m_start"):
        FreeParam("m", "m_start", None, 1.0, 0.0, 1.0, 0.5, mode="add")


def _erfc_cheb(z):
    """The MLP kernels' erfc (csrc/bmc_mlp.hip erfc_cheb), the Chebyshev fit of Numerical Recipes' erfcc,
    evaluated in float32 as the kernels do."""
    c = [0.17087277, -0.82215223, 1.48851587, -1.13520398, 0.27886807, -0.18628806, 0.09678418, 0.37409196,
         1.00002368, -1.26551223]
    z = np.asarray(z, np.float32)
    a = np.abs(z)
    t = np.float32(1) / (np.float32(1) + np.float32(0.5) * a)
    p = np.float32(c[0])
    for ck in c[1:]:
        p = t * p + np.float32(ck)
    e = t * np.exp(-a * a + p)
    return np.where(z >= 0, e, np.float32(2) - e)


def test_the_mlp_gelu_fit_against_scipy():
    """GELU x erfc(-x / sqrt 2) / 2 with the fit's erfc against scipy's erfc: far below BF16's resolution
    (the activations are rounded to BF16, 3.9e-3 relative), and a dropped coefficient is not."""
    from scipy.special import erfc
    x = np.linspace(-12.0, 12.0, 400001)
    ref = 0.5 * x * erfc(-x / np.sqrt(2.0))
    got = 0.5 * x * _erfc_cheb(-x / np.sqrt(2.0)).astype(np.float64)
    assert np.max(np.abs(got - ref)) < 3e-7
    big = np.abs(ref) > 1e-6
    assert np.max(np.abs(got - ref)[big] / np.abs(ref[big])) < 5e-6
    z = np.linspace(0.0, 6.0, 60001)
    assert np.max(np.abs(_erfc_cheb(z) - erfc(z)) / erfc(z)) < 1e-5
    # the plausible mistake of dropping the last coefficient of P
    c0 = _erfc_cheb(z)
    t = 1.0 / (1.0 + 0.5 * z)
    assert np.max(np.abs(c0 / np.exp(0.17087277 * t ** 9) - erfc(z)) / erfc(z)) > 1e-2
