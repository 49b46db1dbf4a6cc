"""Pins the CPU oracle (oracle/) before anything is checked against it. CPU only.

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a wrong sign or
index, a transposed operand) fails: closed forms of the Bloch-McConnell equations, scipy's
exponential and Frechet derivative, two independent Jacobian routes, textbook lineshapes, the
reference's own invariants (tests/test_solver.py, test_config.py) and the LM rules of lm.py.
"""
import numpy as np
import pytest
import scipy.linalg

import bmc_oracle as O
import oracle_c as OC
from helpers import cfg1_events, cfg1_params, four_pool_params, small_train

ROW = lambda ev, o: {k: np.asarray(v)[o] for k, v in ev.items()}


def delay_row(dt, freq=0.0):
    return dict(type=np.array([0]), dt=np.array([dt]), rf_amp=np.array([0.0]), rf_phase=np.array([0.0]),
                rf_freq=np.array([freq]))


def test_pure_relaxation_closed_form():
    """No RF, no pools: Mz = 1 + (Mz0 - 1) exp(-R1 t); transverse decays as exp(-R2 t) and precesses at off_w."""
    p = O.make_params(1.3, 0.07, pools=[dict(f=0.0, k=0.0, dw=3.5, t1=1.0, t2=0.1)])
    M0 = O.initial_state(p) * 0.3
    M0[0], M0[2] = 0.2, -0.1     # water x, y (nt = 2 -> y index 2)
    t, f = 0.7, 13.0
    M = O.simulate_sequence(p, delay_row(t, f), M0, return_trajectory=False)
    assert M[4] == pytest.approx(1 + (M0[4] - 1) * np.exp(-t / 1.3), rel=1e-13)
    w = 2 * np.pi * f
    c, s = np.cos(w * t), np.sin(w * t)
    e2 = np.exp(-t / 0.07)
    # dx/dt = -R2 x + w y, dy/dt = -w x - R2 y  (matrix.py:238-246)
    assert M[0] == pytest.approx(e2 * (c * M0[0] + s * M0[2]), rel=1e-12)
    assert M[2] == pytest.approx(e2 * (-s * M0[0] + c * M0[2]), rel=1e-12)


def test_cw_steady_state_is_minus_A_inverse_C():
    """A long CW block reaches the fixed point of dM/dt = A M + C."""
    p = cfg1_params()
    ev = O.cw_saturation([2.0], b1_uT=1.0, tsat_s=60.0, b0_T=3.0)
    M = O.simulate_sequence(p, ROW(ev, 0), O.initial_state(p))
    A, C = O.bmc_matrix(p, float(ev["rf_amp"][0, 0]), 0.0, float(ev["rf_freq"][0, 0]))
    keep = np.arange(A.shape[0] - 1)          # the empty MT slot (no MT pool) is a zero row of A
    ss = np.linalg.solve(A[np.ix_(keep, keep)], -C[keep])
    assert np.max(np.abs(M[keep] - ss)) < 1e-12


def test_zero_duration_events_are_identities():
    p = four_pool_params()
    M0 = O.initial_state(p) * 0.9
    row = dict(type=np.array([0, 1, 0]), dt=np.array([0.0, 1e-10, 0.0]), rf_amp=np.array([0.0, 100.0, 0.0]),
               rf_phase=np.zeros(3), rf_freq=np.array([0.0, 300.0, 0.0]))
    assert np.array_equal(O.simulate_sequence(p, row, M0), M0)


def test_far_wing_reads_water_f():
    """At -300 ppm with B1 = 0, water Mz reads water_f = 1 (reference tests/test_config.py:118-134); and the
    water relaxation target is 1, not water_f (C[zw] = r1_w * 1.0, matrix.py:289-290)."""
    p = O.make_params(1.3, 0.05, pools=[dict(f=0.0009, k=50.0, dw=3.5, t1=1.0, t2=0.1)])
    ev = O.cw_saturation([-300.0], b1_uT=0.0, tsat_s=2.0, b0_T=3.0)
    M = O.simulate_sequence(p, ROW(ev, 0), O.initial_state(p))
    assert M[O.mz_loc(p)] == pytest.approx(1.0, rel=1e-10)     # 16 squarings of a 2.4e5 rad/s rotation
    q = O.make_params(1.3, 0.05, pools=[dict(f=0.0, k=0.0, dw=3.5, t1=1.0, t2=0.1)], water_f=0.9)
    M = O.simulate_sequence(q, ROW(ev, 0), O.initial_state(q))
    assert M[O.mz_loc(q)] == pytest.approx(1 + (0.9 - 1) * np.exp(-2.0 / 1.3), rel=1e-12)


def test_lorentzian_and_superlorentzian():
    assert O.lorentzian(0.0, 0.02) == 0.02
    assert O.lorentzian(100.0, 0.01) == pytest.approx(0.01 / 2.0)
    # SuperLorentzian: an independent restatement of the 101-point rectangle rule (matrix.py:26-38)
    t2, d = 1e-5, 4000.0
    acc = 0.0
    for i in range(101):
        u = i / 100.0
        T = max(abs(3 * u * u - 1), 1e-6)
        acc += np.sqrt(2 / np.pi) * t2 / T * np.exp(-2 * (d * t2 / T) ** 2)
    assert O.interpolate_sl(d, t2) == pytest.approx(acc * np.pi * 0.01, rel=1e-13)
    # the cubic bridge meets the SuperLorentzian at its nodes (c = 0 at px1 - 1, c = 1 at px2)
    w0 = 800.0
    assert O.interpolate_chs(-100 - w0 - 1, t2, w0) == pytest.approx(O.interpolate_sl(-100 - w0, t2), rel=1e-12)
    assert O.interpolate_chs(100 + w0, t2, w0) == pytest.approx(O.interpolate_sl(100 + w0, t2), rel=1e-12)


def test_c_expm_matches_scipy_and_frechet():
    rng = np.random.default_rng(7)
    for n, s in [(8, 0.5), (11, 3.0), (17, 40.0)]:
        X = rng.normal(size=(n, n)) * s / n
        dX = rng.normal(size=(2, n, n))
        E, dE = OC.expm(X, dX)
        assert np.max(np.abs(E - scipy.linalg.expm(X))) < 1e-13 * max(1.0, np.abs(E).max())
        for k in range(2):
            L = scipy.linalg.expm_frechet(X, dX[k], compute_expm=False)
            assert np.max(np.abs(dE[k] - L)) < 1e-11 * max(1.0, np.abs(L).max())


def test_c_oracle_matches_numpy_oracle():
    p = four_pool_params(b0i=0.1)
    ev = small_train(np.array([-3.0, 0.5, 3.5]), n_pulses=2, n_samples=6)
    m_c, _ = OC.simulate(OC.pack(p)[None], 2, ev, O.initial_state(p))
    m_n = O.simulate_spectrum(p, ev, O.initial_state(p))
    assert np.max(np.abs(m_c[0] - m_n)) < 1e-12


def test_two_jacobian_routes_agree():
    """Central differences of Z vs scipy's exact Frechet derivative chained through the scan."""
    p = cfg1_params()
    ev = {k: v[::10] for k, v in cfg1_events().items()}
    slots = [("cest_k", 0), ("cest_f", 0), ("water_r2", None)]
    x = np.array([50.0, 9.01e-4, 1 / 0.075])
    m0 = O.initial_state(p)
    J1 = O.jacobian_fd(p, ev, m0, O.mz_loc(p), slots, x)
    J2 = O.jacobian_frechet(p, ev, m0, O.mz_loc(p), slots, x)
    for c in range(3):
        assert np.max(np.abs(J1[:, c] - J2[:, c])) < 1e-6 * np.abs(J2[:, c]).max()


def test_c_forward_mode_jacobian_matches_frechet_route():
    p = four_pool_params()
    ev = small_train(np.array([-2.5, 3.5]), n_pulses=2, n_samples=5)
    slots = [("cest_k", 0), ("b0_inhom", None), ("rel_b1", None)]
    x = np.array([50.0, 0.0, 1.0])
    idx = [OC.slot_index(f, i, 2) for f, i in slots]
    _, dm = OC.simulate(OC.pack(p)[None], 2, ev, O.initial_state(p), slots=idx)
    J = O.jacobian_frechet(p, ev, O.initial_state(p), O.mz_loc(p), slots, x)
    for c in range(3):
        assert np.max(np.abs(dm[0, :, c, 6] - J[:, c])) < 1e-7 * max(np.abs(J[:, c]).max(), 1e-12)


def test_reference_float32_algorithm_has_its_documented_floor():
    """The reference's own float32 expm errs by ~1e-4 on configuration 1 (settings.py:10-11)."""
    p = cfg1_params()
    ev = cfg1_events()
    z64 = O.simulate_spectrum(p, ev, O.initial_state(p))[:, 4]
    z32 = O.simulate_spectrum(p, ev, O.initial_state(p), expm=O.expm_ref32)[:, 4]
    err = np.max(np.abs(z32 - z64))
    assert 1e-5 < err < 1e-3


def test_physics_invariants_on_the_oracle():
    """tests/test_solver.py of the reference, on the oracle: far wing, symmetry without a pool, dip side, k = 0."""
    ppm = np.linspace(-6, 6, 121)
    ev = O.cw_saturation(ppm, b1_uT=1.0, tsat_s=2.0, b0_T=2.8936)

    def z(f=0.005, k=300.0):
        p = O.make_params(3.0, 0.6, pools=[dict(f=f, k=k, dw=3.5, t1=1.0, t2=0.01)], b0=2.8936)
        return O.simulate_spectrum(p, ev, O.initial_state(p))[:, 4]
    base = z()
    assert 0.95 < base[0] < 1.01
    sym = z(f=0.0)
    assert np.abs(sym - sym[::-1]).max() < 2e-3
    i_pos = int(np.argmin(np.abs(ppm - 3.5)))
    assert base[120 - i_pos] - base[i_pos] > 0.05
    k0 = z(k=0.0)
    assert abs(k0[120 - i_pos] - k0[i_pos]) < 2e-3


def test_lm_rosenbrock_and_held_entry():
    """lm.py restated: Rosenbrock to (1, 1) (test_fitting.py:90-100); a held entry never moves."""
    r = lambda x: np.array([10.0 * (x[1] - x[0] ** 2), 1.0 - x[0]])
    J = lambda x: np.array([[-20.0 * x[0], 10.0], [-1.0, 0.0]])
    x = O.levenberg_marquardt(r, J, [-1.2, 1.0], [-5, -5], [5, 5], n_iter=200)
    assert np.allclose(x, [1.0, 1.0], atol=1e-3)
    r2 = lambda x: np.array([x[0] - 1.0, x[1] - 2.0, 0.1 * x[0] * x[1]])
    J2 = lambda x: np.array([[1.0, 0.0], [0.0, 1.0], [0.1 * x[1], 0.1 * x[0]]])
    held = O.levenberg_marquardt(r2, J2, [0.0, 5.0], [-10, -10], [10, 10], n_iter=30, free=np.array([1.0, 0.0]))
    assert held[1] == 5.0


def test_lm_trial_active_bounds_is_the_reduced_problem():
    """joint.py:149-162 restated: a railed value with an outward step is held (test_joint.py:125-143)."""
    rng = np.random.default_rng(1)
    J = rng.normal(size=(30, 3))
    r = rng.normal(size=30)
    if (J.T @ r)[0] < 0:
        J[:, 0] *= -1
    x = np.array([0.0, 0.5, 0.5])
    lo, hi = np.zeros(3), np.ones(3) * 10
    held = O.lm_trial(x, 0.1, J, r, lo, hi, active_bounds=True)
    Jr = J[:, 1:]
    H, g = Jr.T @ Jr, Jr.T @ r
    want = x[1:] + np.linalg.solve(H + 0.1 * (np.diag(np.diag(H)) + 1e-9 * np.eye(2)), -g)
    assert held[0] == 0.0
    np.testing.assert_allclose(held[1:], want, rtol=1e-12)


def test_golden_configuration_1_is_reproduced():
    """The committed float64 fixture of configuration 1 (tests/golden/make_golden.py) still holds."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1_f64.npz"))
    p = cfg1_params()
    ev = cfg1_events()
    idx = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1)]
    m, dm = OC.simulate(OC.pack(p)[None], 1, ev, O.initial_state(p), slots=idx)
    np.testing.assert_allclose(m[0, :, 4], g["z"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(dm[0, :, :, 4], g["jac"], rtol=1e-10, atol=0)


# ---- pins added for the lineshapes, the MT block, the LM rules and the CW amplitude -----------------

def _integral(f, edges):
    from scipy.integrate import quad
    return sum(quad(f, a, b, limit=400, epsabs=0.0, epsrel=1e-11)[0] for a, b in zip(edges[:-1], edges[1:]))


def test_lineshape_integrals():
    """Every node of the SuperLorentzian rule is a unit-area Gaussian in the offset, so the 101-node rule
    with its pi * du weight integrates to exactly 1.01 pi; the Lorentzian t2 / (1 + (d t2)^2) carries no
    1/pi and integrates to pi (matrix.py:26-38, 70-81). A dropped pi, du or sqrt(2/pi), or a lineshape
    normalised to 1, fails by a factor."""
    t2 = 1e-5
    # node widths run from 0.0092 / (2 t2) to 2 / (2 t2) rad/s: split the half line where they change
    edges = [0.0, 2e3, 1e4, 5e4, 2e5, 6e5, 3e6]
    half = _integral(lambda d: O.interpolate_sl(d, t2), edges)
    assert 2 * half == pytest.approx(1.01 * np.pi, rel=1e-8)
    t2l = 0.01
    from scipy.integrate import quad
    half_l = quad(lambda d: O.lorentzian(d, t2l), 0.0, np.inf, epsabs=0.0, epsrel=1e-12)[0]
    assert 2 * half_l == pytest.approx(np.pi, rel=1e-9)


def test_mt_saturation_rate_is_henkelman():
    """The MT pool's RF saturation rate is the textbook pi w1^2 g(delta) with g a unit-area lineshape
    (Henkelman et al. 1993): for the Lorentzian, g = (T2 / pi) / (1 + (delta T2)^2). The diagonal of the
    MT row must be -(R1 + k + pi w1^2 g), with w1 = 2 pi amp rel_b1 (matrix.py:121, 210, 282-286)."""
    p = O.make_params(1.3, 0.07, pools=[dict(f=2e-3, k=100.0, dw=3.5, t1=1.0, t2=0.02)],
                      mt=dict(f=0.12, k=25.0, dw=-2.5, t1=1.1, t2=1e-5, lineshape="lorentzian"), rel_b1=0.9)
    amp, freq = 85.0, 1500.0
    A, _ = O.bmc_matrix(p, amp, 0.3, freq)
    larmor = p["gamma"] * 1e6 * p["b0"]
    delta = 2 * np.pi * freq - p["mt_dw"] * 1e-6 * larmor          # MT offset: RF frequency past the MT line
    w1 = 2 * np.pi * amp * p["rel_b1"]
    t2 = 1.0 / p["mt_r2"]
    g = (t2 / np.pi) / (1 + (delta * t2) ** 2)
    zmt = 3 * (1 + 1)
    assert A[zmt, zmt] == pytest.approx(-(p["mt_r1"] + p["mt_k"] + np.pi * w1 ** 2 * g), rel=1e-13)


@pytest.mark.parametrize("lineshape", ["lorentzian", "superlorentzian"])
def test_equilibrium_with_an_mt_pool_is_stationary(lineshape):
    """With no RF, the start state (water 1, pools f, MT f_mt) is the equilibrium of the exchange: it must
    not move over 10 s in either oracle. Fails if C[zmt] lacks f_mt, k_mt is not weighted by f_mt on the
    water side, or a sign or index of the MT block is wrong (matrix.py:282-293)."""
    p = O.make_params(1.3, 0.07, pools=[dict(f=3e-3, k=200.0, dw=3.5, t1=1.0, t2=0.02),
                                        dict(f=6e-3, k=900.0, dw=2.0, t1=1.2, t2=0.03)],
                      mt=dict(f=0.13, k=40.0, dw=-2.5, t1=0.9, t2=1e-5, lineshape=lineshape))
    M0 = O.initial_state(p)
    A, C = O.bmc_matrix(p, 0.0, 0.0, 250.0)
    assert np.max(np.abs(A @ M0 + C)) < 1e-14
    row = dict(type=np.array([0, 0]), dt=np.array([4.0, 6.0]), rf_amp=np.zeros(2), rf_phase=np.zeros(2),
               rf_freq=np.array([0.0, 250.0]))
    M = O.simulate_sequence(p, row, M0)
    assert np.max(np.abs(M - M0)) < 1e-12
    ev = {k: v[None] for k, v in row.items()}
    m_c, _ = OC.simulate(OC.pack(p)[None], 2, ev, M0)
    assert np.max(np.abs(m_c[0, 0] - M0)) < 1e-12
    # the check bites: the MT source unweighted by its fraction, or the water-side rate unweighted, move it
    zmt, zw = 3 * 3, 2 * 3
    bad_c = C.copy()
    bad_c[zmt] = p["mt_r1"]
    bad_a = A.copy()
    bad_a[zw, zw] += p["mt_k"] * p["mt_f"] - p["mt_k"]
    assert np.max(np.abs(A @ M0 + bad_c)) > 1e-3 and np.max(np.abs(bad_a @ M0 + C)) > 1e-3


def test_cw_nutation_closed_form():
    """Relaxation and pools removed, water alone under CW saturation from Mz = 1 nutates about the
    effective field: Mz(t) = (d^2 + w1^2 cos(W t)) / W^2 with W^2 = d^2 + w1^2, w1 = gamma B1 =
    b1 * 267.5153 rad/s and d = 2 pi ppm 42.5764 B0 (saturation.py:79-83, matrix.py:121-162). The float32
    storage of the event values (saturation.py:79-83) bounds the agreement to ~1e-6."""
    p = O.make_params(np.inf, np.inf, pools=[], b0=3.0)
    assert p["water_r1"] == 0.0 and p["water_r2"] == 0.0
    b1, t = 1.3, 0.05
    ppm = np.array([0.0, 0.002, -0.004])
    ev = O.cw_saturation(ppm, b1_uT=b1, tsat_s=t, b0_T=3.0)
    Mz = O.simulate_spectrum(p, ev, O.initial_state(p))[:, 2]
    w1 = b1 * 267.5153
    d = 2 * np.pi * ppm * 42.5764 * 3.0
    W2 = d ** 2 + w1 ** 2
    want = (d ** 2 + w1 ** 2 * np.cos(np.sqrt(W2) * t)) / W2
    np.testing.assert_allclose(Mz, want, rtol=0, atol=5e-6)
    # gamma in rad/s/uT and in Hz/T are the same constant
    assert 267.5153 / (2 * np.pi) == pytest.approx(42.5764, rel=2e-6)


def test_lm_trajectory_by_hand():
    """lm.py:43-61 on r(x) = atan(x - 3), x0 = 0, lam0 = 1e-2, no bounds in play. By hand, with
    J = 1/10, r = atan(-3): delta = -J r / (J^2 (1 + lam) + 1e-9 lam). The first five trials overshoot
    past x = 6.9 and raise the cost (each rejected, lam x 3: 0.01, 0.03, 0.09, 0.27, 0.81); the sixth, at
    lam = 2.43, lands at x = 3.6415326061487 and is accepted (lam x 0.5); then 3.27803114950289 and
    3.09629255683953. A swapped factor, a missing diag(H) or a rejected-step update of x fails."""
    r = lambda x: np.array([np.arctan(x[0] - 3.0)])
    J = lambda x: np.array([[1.0 / (1.0 + (x[0] - 3.0) ** 2)]])
    x, X, C = O.levenberg_marquardt(r, J, [0.0], [-100.0], [100.0], n_iter=8, trajectory=True)
    np.testing.assert_array_equal(X[:5, 0], 0.0)
    np.testing.assert_allclose(C[:5], np.arctan(3.0) ** 2, rtol=1e-15)
    np.testing.assert_allclose(X[5:, 0], [3.6415326061487225, 3.2780311495028904, 3.0962925568395328], rtol=1e-13)


def test_lm_damping_floor_and_strict_accept():
    """The 1e-9 floor is scaled by lam, not added to diag(H) alone: with J = 1e-6 the first step of
    r(x) = 1e-6 x - 1 from 0 is 1e-6 / (1e-12 * 1.01 + 1e-11) = 90826.5 (without the floor 990099, with an
    unscaled floor 998.99). A clipped step that leaves x where it was ties the cost and is rejected
    (strict <): lam grows x 3 every iteration, x never moves."""
    r = lambda x: np.array([1e-6 * x[0] - 1.0])
    J = lambda x: np.array([[1e-6]])
    x, X, _ = O.levenberg_marquardt(r, J, [0.0], [-1e7], [1e7], n_iter=1, trajectory=True)
    assert X[0, 0] == pytest.approx(90826.5213442325, rel=1e-12)
    r2 = lambda x: np.array([x[0] - 3.0])
    J2 = lambda x: np.array([[1.0]])
    x, X, C = O.levenberg_marquardt(r2, J2, [0.0], [-1.0], [0.0], n_iter=4, trajectory=True)
    np.testing.assert_array_equal(X[:, 0], 0.0)
    np.testing.assert_array_equal(C, 9.0)
    # lam after 4 ties is 0.01 * 3^4; the next free step from 0 towards 3 shows it
    x5, X5, _ = O.levenberg_marquardt(r2, J2, [0.0], [-1.0], [10.0], n_iter=1, lam0=0.01 * 3 ** 4, trajectory=True)
    assert X5[0, 0] == pytest.approx(3.0 / (1.0 + 0.81 * (1 + 1e-9)), rel=1e-14)
