"""bmc_expm, the dense exponential behind expm_library / expm_plain / propagate
(src/bmc_jax/simulation/expm.py:26-119 of the reference), on the GPU through the C ABI.

float64 against scipy's expm (Al-Mohy--Higham, an independent algorithm); float32 against the oracle's
restatement of the reference's float32 algorithm (Pade 3/5/7, Gauss-Jordan, expm.py:50-83) and its NaN
rule. The matrices are Bloch-McConnell generators scaled so that every Pade order and squaring count of
both rules is reached."""
import numpy as np
import pytest
import scipy.linalg

import bmc_oracle as O

pytestmark = pytest.mark.gpu

# 1-norms spanning the float64 bands (orders 3, 5, 7, 9, 13 and squarings) and the float32 ones
NORMS = [0.0, 1e-3, 0.1, 0.4, 0.7, 1.5, 2.0, 3.0, 5.0, 11.0, 90.0, 3000.0, 2.0e5]


def generators(seed=0):
    """Augmented Bloch-McConnell generators [[A, C], [0, 0]] of two CEST pools + MT (size 11), scaled to NORMS."""
    rng = np.random.default_rng(seed)
    out = []
    for target in NORMS:
        p = O.make_params(rng.uniform(0.8, 2.0), rng.uniform(0.05, 0.2),
                          pools=[dict(f=rng.uniform(1e-4, 1e-2), k=rng.uniform(20, 3000), dw=rng.uniform(-4, 4),
                                      t1=1.0, t2=rng.uniform(0.005, 0.1)) for _ in range(2)],
                          mt=dict(f=0.1, k=30.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape="Lorentzian"), b0=3.0)
        A, C = O.bmc_matrix(p, rng.uniform(0, 300), rng.uniform(0, 6), rng.uniform(-2000, 2000))
        X = O.augmented(A, C, 1.0)
        l1 = np.max(np.sum(np.abs(X), axis=0))
        out.append(X * (target / l1 if target > 0 else 0.0))
    return np.stack(out)


@pytest.fixture(scope="module")
def nat():
    from bmc_jax import _native
    _native.torch_cuda()
    return _native


def _l1(x):
    return float(np.max(np.sum(np.abs(np.asarray(x, np.float64)), axis=0)))


def test_float64_matches_scipy(nat):
    """Each squaring doubles the rounding error: below 9 squarings the two algorithms agree to 2e-13;
    at 15 (norm 2e5) the Pade rule itself sits 3.6e-12 from scipy (measured on a NumPy restatement)."""
    X = generators()
    E = nat.expm(X)
    for x, e in zip(X, E):
        ref = scipy.linalg.expm(x)
        scale = max(1.0, np.abs(ref).max())
        tol = 2e-13 if _l1(x) < 1000 else 2e-11
        assert np.abs(e - ref).max() / scale < tol, _l1(x)
    np.testing.assert_array_equal(E[0], np.eye(X.shape[1]))


def test_float32_matches_the_reference_rule(nat):
    X = generators(1).astype(np.float32)
    E = nat.expm(X)
    for x, e in zip(X, E):
        ref = O.expm_ref32(x)
        scale = max(1.0, np.abs(ref).max())
        # float32 rounding (6e-8) doubles with each squaring: 9 at norm 3000, 15 at norm 2e5
        tol = 2e-6 if _l1(x) < 100 else 2e-4
        assert np.abs(e - ref).max() / scale < tol, _l1(x)
        exact = scipy.linalg.expm(x.astype(np.float64))
        assert np.abs(e - exact).max() / max(1.0, np.abs(exact).max()) < tol * 2, _l1(x)


def test_nan_past_sixteen_squarings(nat):
    """2^17 times theta_m: 17 squarings by the floor rule, more than the 16 the reference allows."""
    for dtype, theta in ((np.float32, 3.925724783138660), (np.float64, 5.371920351148152)):
        x = generators(2)[3].astype(dtype)
        x = x / np.max(np.sum(np.abs(x), axis=0)) * theta * 2.0 ** 17 * 1.01
        assert np.isnan(nat.expm(x)).all()
        ok = x / 4.0                                     # 15 squarings: finite
        assert np.isfinite(nat.expm(ok)).all()
    assert np.isnan(O.expm_ref32(generators(2)[3].astype(np.float32) * 1e7)).all()


def test_batch_axes_and_sizes(nat):
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 17, 24):
        a = rng.standard_normal((3, 2, n, n)) * 0.7
        e = nat.expm(a)
        assert e.shape == a.shape
        for idx in np.ndindex(3, 2):
            np.testing.assert_allclose(e[idx], scipy.linalg.expm(a[idx]), rtol=1e-12, atol=1e-13)
    with pytest.raises(Exception):
        nat.expm(np.zeros((25, 25)))


def test_propagate_both_backends(nat):
    from bmc_jax import settings
    from bmc_jax.simulation import expm_plain, propagate
    p = O.make_params(1.3, 0.075, pools=[dict(f=9.01e-4, k=50.0, dw=3.5, t1=1.3, t2=0.1)], b0=3.0)
    A, C = O.bmc_matrix(p, 42.0, 0.3, 150.0)
    M = O.initial_state(p)
    want = (scipy.linalg.expm(O.augmented(A, C, 0.05)) @ np.append(M, 1.0))[:-1]
    np.testing.assert_allclose(propagate(M, 0.05, A, C), want, rtol=1e-12, atol=1e-14)
    settings.use_expm("plain")
    try:
        got = propagate(M.astype(np.float32), 0.05, A.astype(np.float32), C.astype(np.float32))
        assert got.dtype == np.float32
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
        with pytest.raises(ValueError, match="float32"):
            expm_plain(np.eye(3))
    finally:
        settings.use_expm("library")
