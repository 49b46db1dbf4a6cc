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
