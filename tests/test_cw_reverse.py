"""The reverse accumulation behind the single-event Jacobian kernel (csrc/bmc_cwj.hip), checked on the CPU
with scipy's exponential and Frechet derivative, and (GPU) the kernel against the forward-mode kernel.

For E = T^n, T = exp(Y), Y = X / n, n = 2^s and Z = e_z^T E v, the kernel computes
dZ/dp = <dT_p, G_n> with G_1 = e_z v^T and G_2n = G_n E_k^T + E_k^T G_n (E_k = T^n at that level),
keeping only the rows of G above the augmented one. Here the same recursion in NumPy must give
e_z^T L(X, dX_p) v, L scipy's Frechet derivative of exp at the unscaled X (an independent route)."""

import numpy as np
import pytest
import scipy.linalg as sl

import bmc_oracle as O
import oracle_c as OC
from helpers import four_pool_params, pack_many


def _generator(p, amp, ph, fr, dt):
    A, C = O.bmc_matrix(p, amp, ph, fr)
    return O.augmented(A, C, dt)


def _level(E, G, N):
    """One level of the kernel's recursion on (N+1)x(N+1) E and the N x (N+1) rows of G."""
    Gf = np.vstack([G, np.zeros((1, N + 1))])      # the augmented row never feeds the rows above it
    Gn = (Gf @ E.T + E.T @ Gf)[:N]
    return E @ E, Gn


@pytest.mark.parametrize("s", [0, 1, 5, 9])
def test_reverse_accumulation_matches_the_frechet_derivative(s):
    p = four_pool_params(k_am=300.0, f_am=2e-3, b0i=0.05, lineshape="lorentzian")
    nt = 1 + len(p["cest_k"])
    N = 3 * nt + 1
    jz = 2 * nt
    X = _generator(p, 40.0, 0.3, 270.0, 0.37)
    v = np.r_[O.initial_state(p, 0.8), 1.0]
    # directions: water R1 and the amide exchange rate, by central differences of the generator
    dXs = []
    for f, i in (("water_r1", None), ("cest_k", 0)):
        def gen(h, f=f, i=i):
            q = {k: (np.array(val, copy=True) if isinstance(val, np.ndarray) else val) for k, val in p.items()}
            if i is None:
                q[f] = q[f] + h
            else:
                q[f][i] += h
            return _generator(q, 40.0, 0.3, 270.0, 0.37)
        h = 1e-6 * max(1.0, abs(p[f] if i is None else p[f][i]))
        dXs.append((gen(h) - gen(-h)) / (2 * h))
    n = 2 ** s
    Y = X / n
    T = sl.expm(Y)
    E, G = T.copy(), np.zeros((N, N + 1))
    G[jz] = v                                            # G_1 = e_z v^T
    for _ in range(s):
        E, G = _level(E, G, N)
    np.testing.assert_allclose(E, sl.expm(X), rtol=1e-10, atol=1e-12)
    for dX in dXs:
        dT = sl.expm_frechet(Y, dX / n, compute_expm=False)
        got = float(np.sum(dT[:N] * G))
        want = float(sl.expm_frechet(X, dX, compute_expm=False)[jz] @ v)
        assert got == pytest.approx(want, rel=1e-9, abs=1e-12)
    # a plausible mistake, G_2n = E^T G_n E^T (a product instead of the sum), gives something else
    E2, Gb = T.copy(), np.zeros((N, N + 1))
    Gb[jz] = v
    for _ in range(s):
        Gf = np.vstack([Gb, np.zeros((1, N + 1))])
        Gb = (E2.T @ Gf @ E2.T)[:N]
        E2 = E2 @ E2
    if s > 0:
        dT = sl.expm_frechet(Y, dXs[1] / n, compute_expm=False)
        assert abs(float(np.sum(dT[:N] * Gb)) - float(sl.expm_frechet(X, dXs[1], compute_expm=False)[jz] @ v)) > 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [0, 1, 2, 3, 4])
def test_reverse_jacobian_matches_forward_mode(nc):
    """bmc_simulate on CW rows: the reverse kernel (default) against the forward-mode kernel
    (bmc_simulate_ex with BMC_SIM_FORWARD_JACOBIAN), every tangent kind that a CW row can carry, float64."""
    import torch
    from bmc_jax import _native as nat
    nat.torch_cuda()
    pools = [dict(f=2e-3 * (i + 1), k=80.0 * (3 ** i), dw=1.0 + 0.9 * i, t1=1.0, t2=0.02 + 0.01 * i) for i in range(nc)]
    recs = [O.make_params(1.2 + 0.3 * v, 0.07, pools=pools, b0=3.0, b0_inhom=0.03 * v, rel_b1=0.95 + 0.03 * v,
                          mt=(dict(f=0.08, k=25.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape="superlorentzian") if v % 2 else None),
                          n_cest=nc)
            for v in range(3)]
    # one pool model per launch: MT slot present in all (zero where absent)
    ev = O.cw_saturation(np.linspace(-6, 6, 25), b1_uT=1.5, tsat_s=2.0, b0_T=3.0)
    names = [("water_r1", None), ("water_r2", None), ("b0_inhom", None), ("rel_b1", None), ("mt_f", None),
             ("mt_k", None), ("mt_dw", None), ("mt_r1", None), ("mt_r2", None), ("b0", None), ("gamma", None)]
    names += [(f, 0) for f in ("cest_k", "cest_f", "cest_dw", "cest_r1", "cest_r2")] if nc else []
    slots = [OC.slot_index(f, i, nc) for f, i in names]
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    prm = nat.device_array(pack_many(recs), torch.float64)
    m0 = nat.device_array(np.stack([O.initial_state(q) for q in recs]), torch.float64)
    jc = 2 * (1 + nc)
    out = {}
    for mode in ("reverse", "forward"):
        m, J = nat.simulate(plan, prm, nc, nat.BMC_F64, m0, slots=slots, scales=[1.0] * len(slots), jc=jc,
                            forward_jacobian=(mode == "forward"))
        torch.cuda.synchronize()
        out[mode] = (m.cpu().numpy(), J.cpu().numpy())
    (mr, Jr), (mf, Jf) = out["reverse"], out["forward"]
    # the kernels split exp(X) differently (Taylor degree cap 18 with more squarings against 30):
    # the states agree to rounding, within the 1e-11 that both keep against the float64 oracle
    assert np.max(np.abs(mr - mf)) < 1e-11
    for p in range(len(slots)):
        scale = max(np.abs(Jf[..., p]).max(), 1e-30)
        assert np.max(np.abs(Jr[..., p] - Jf[..., p])) / scale < 1e-9, names[p]


def test_single_event_rule():
    """Which event tables take the reverse kernel: CW rows (front padding allowed), not pulsed trains,
    spoilers, or a readout before the row's end."""
    from bmc_jax import _flops
    from bmc_jax.sequences import cw_saturation, pulsed_train
    from bmc_jax.simulation.events import pad_events
    ev = pad_events(cw_saturation(np.linspace(-3, 3, 5), 1.0, 2.0, 3.0), 4)   # [pad, pad, RF, ADC]
    np.testing.assert_array_equal(_flops.single_event_rows(ev), [2] * 5)
    assert _flops.single_event_rows(pulsed_train(np.array([1.0]), n_pulses=2, n_samples=3)) is None
    t = np.array(ev.type, copy=True)
    t[0, 0] = 4                                    # a zero-length spoiler still zeroes the transverse state
    assert _flops.single_event_rows(ev._replace(type=t)) is None
    t = np.array(ev.type, copy=True)
    t[1, 3] = 0                                    # the readout no longer last ...
    t[1, 1] = 3                                    # ... and one before the block: still fine (reset to M0)
    d = np.array(ev.dt, copy=True)
    assert _flops.single_event_rows(ev._replace(type=t, dt=d)) is not None
    t2 = np.array(ev.type, copy=True)
    t2 = np.concatenate([t2, np.zeros((5, 1), t2.dtype)], axis=1)   # an event after the readout
    ev2 = ev._replace(type=t2, dt=np.concatenate([d, np.zeros((5, 1))], axis=1))
    assert _flops.single_event_rows(ev2) is None


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [1, 4])
def test_reverse_kernel_edge_rows(nc):
    """Rows with no propagating event (Z = M0[jc], zero Jacobian), a row past the reference's NaN rule
    (squaring count above 16: NaN state and Jacobian) and ordinary CW rows, reverse against forward;
    at four CEST pools the kernel spreads the augmented column over its 16-lane team."""
    import torch
    from bmc_jax import _native as nat
    nat.torch_cuda()
    pools = [dict(f=9.01e-4 * (i + 1), k=50.0 * (2 ** i), dw=3.5 - 1.1 * i, t1=1.3, t2=0.1) for i in range(nc)]
    p = O.make_params(1.3, 0.075, pools=pools, b0=3.0, n_cest=nc)
    ev = O.cw_saturation(np.linspace(-6, 6, 6), b1_uT=1.0, tsat_s=2.0, b0_T=3.0)
    ev = {k: np.array(v, copy=True) for k, v in ev.items()}
    ev["dt"][1, 0] = 0.0            # no propagating event
    ev["dt"][2, 0] = 5e3            # 1-norm far beyond theta_ref 2^17: the reference gives NaN
    slots = [OC.slot_index("cest_k", 0, nc), OC.slot_index("cest_f", 0, nc), OC.slot_index("water_r2", None, nc)]
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    recs = [p, p]
    prm = nat.device_array(pack_many(recs), torch.float64)
    m0s = np.stack([O.initial_state(q) for q in recs]) * np.array([[1.0], [0.6]])
    m0 = nat.device_array(m0s, torch.float64)
    out = {}
    for mode in ("reverse", "forward"):
        m, J = nat.simulate(plan, prm, nc, nat.BMC_F64, m0, slots=slots, scales=[1.0] * 3, jc=2 * (1 + nc),
                            forward_jacobian=(mode == "forward"))
        torch.cuda.synchronize()
        out[mode] = (m.cpu().numpy(), J.cpu().numpy())
    (mr, Jr), (mf, Jf) = out["reverse"], out["forward"]
    np.testing.assert_array_equal(np.isnan(mr), np.isnan(mf))
    np.testing.assert_array_equal(np.isnan(Jr), np.isnan(Jf))
    assert np.all(np.isnan(mr[:, 2])) and np.all(np.isnan(Jr[:, 2]))
    np.testing.assert_array_equal(mr[:, 1], m0s)           # nothing propagates: the start state
    assert np.all(Jr[:, 1] == 0.0)
    ok = [0, 3, 4, 5]
    assert np.max(np.abs(mr[:, ok] - mf[:, ok])) < 1e-11
    for c in range(len(slots)):  # each column against its own scale (dZ/dk is far smaller than dZ/df)
        scale = np.abs(Jf[:, ok, c]).max()
        assert scale > 0.0
        assert np.max(np.abs(Jr[:, ok, c] - Jf[:, ok, c])) < 1e-9 * scale, c
