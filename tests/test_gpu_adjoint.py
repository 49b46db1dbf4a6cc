"""The adjoint kernel (bmc_simulate_adjoint, through the C ABI): dL/d rf_amp and dL/d rf_freq for
L = sum gz * Z + sum gj * J.

Pinned three ways:
  * directional derivatives against Richardson-extrapolated central differences of the float64 C
    oracle (one Pade-13 exponential per event; its generator derivative in k and f is exact, the
    generator being linear in both, so the oracle's J carries no differencing noise);
  * single components against central differences of the kernel's own float64 forward;
  * structure: linearity in the cotangents, zero gradient before an ADC reset, float32 I/O.
"""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import pack_many

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nat():
    from bmc_jax import _native
    _native.torch_cuda()
    return _native


def three_pool(k=60.0, f=2e-3, t1w=1.2, t2w=0.06, lineshape="superlorentzian", b0i=0.0):
    """Water, amide +3.5 ppm, semi-solid MT at -2.5 ppm: one CEST pool (state length 7)."""
    return O.make_params(t1w, t2w, pools=[dict(f=f, k=k, dw=3.5, t1=1.3, t2=0.1)],
                         mt=dict(f=0.08, k=25.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape=lineshape), b0=3.0, b0_inhom=b0i)


def two_cest(k=60.0, f=2e-3):
    return O.make_params(1.2, 0.06, pools=[dict(f=f, k=k, dw=3.5, t1=1.3, t2=0.1),
                                           dict(f=4e-3, k=700.0, dw=2.0, t1=1.3, t2=0.05)], b0=3.0)


def water_only():
    return O.make_params(1.2, 0.06, pools=[], b0=3.0)


def three_cest(k=60.0, f=2e-3):
    return O.make_params(1.2, 0.06, pools=[dict(f=f, k=k, dw=3.5, t1=1.3, t2=0.1),
                                           dict(f=4e-3, k=700.0, dw=2.0, t1=1.3, t2=0.05),
                                           dict(f=1e-3, k=150.0, dw=-1.0, t1=1.2, t2=0.02)],
                         mt=dict(f=0.06, k=20.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape="lorentzian"), b0=3.0)


def four_cest(k=60.0, f=2e-3):
    return O.make_params(1.2, 0.06, pools=[dict(f=f, k=k, dw=3.5, t1=1.3, t2=0.1),
                                           dict(f=4e-3, k=700.0, dw=2.0, t1=1.3, t2=0.05),
                                           dict(f=1e-3, k=150.0, dw=-1.0, t1=1.2, t2=0.02),
                                           dict(f=2e-3, k=2000.0, dw=1.2, t1=1.0, t2=0.01)], b0=3.0)


def train(ppm, n_pulses=3, n_samples=12, recovery_s=0.3):
    ev = O.pulsed_train(ppm, b1_peak_uT=2.0, n_pulses=n_pulses, pulse_s=0.01, n_samples=n_samples, gap_s=0.005,
                        recovery_s=recovery_s, b0_T=3.0, dtype=np.float64)
    ev = {k: np.array(v) for k, v in ev.items()}
    # the delays carry the row's frequency too, so every event's frequency matters
    ev["rf_freq"] = np.broadcast_to(ev["rf_freq"].max(axis=1, keepdims=True) + ev["rf_freq"].min(axis=1, keepdims=True),
                                    ev["rf_freq"].shape).copy()
    ev["rf_phase"] = np.tile(np.linspace(0.0, 1.0, ev["dt"].shape[1]), (ev["dt"].shape[0], 1))
    return ev


def slots_for(nc, P):
    s = [OC.slot_index("cest_k", 0, nc), OC.slot_index("cest_f", 0, nc)] if nc else \
        [OC.slot_index("water_r1", None, nc), OC.slot_index("water_r2", None, nc)]
    if P > 2:   # more tangents than one launch carries: water rates, then the last pool's k and f
        s = s + [OC.slot_index("water_r2", None, nc), OC.slot_index("water_r1", None, nc)]
        if nc > 1:
            s += [OC.slot_index("cest_k", nc - 1, nc), OC.slot_index("cest_f", nc - 1, nc)]
    return s[:P]


def scales_for(nc, P):
    base = [100.0, 0.001] if nc else [1.0, 10.0]
    return (base + [10.0, 1.0, 100.0, 0.001])[:P]


def gpu_adjoint(nat, records, ev, slots, scales, gz, gj, dtype=np.float64, want_freq=True):
    import torch
    nc = len(records[0]["cest_k"])
    code = nat.BMC_F64 if dtype == np.float64 else nat.BMC_F32
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    prm = nat.device_array(pack_many(records).astype(dtype), tdt)
    m0 = nat.device_array(np.stack([O.initial_state(p) for p in records]).astype(dtype), tdt)
    ga, gf = nat.simulate_adjoint(plan, prm, nc, code, m0, slots=slots, scales=scales, jc=2 * (1 + nc),
                                  z_cot=torch.as_tensor(gz), jac_cot=None if gj is None else torch.as_tensor(gj),
                                  want_amp=True, want_freq=want_freq)
    torch.cuda.synchronize()
    return ga.cpu().numpy(), (gf.cpu().numpy() if gf is not None else None)


def gpu_loss(nat, records, ev, slots, scales, gz, gj):
    """L from the kernel's float64 forward (bmc_simulate)."""
    import torch
    nc = len(records[0]["cest_k"])
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    prm = nat.device_array(pack_many(records), torch.float64)
    m0 = nat.device_array(np.stack([O.initial_state(p) for p in records]), torch.float64)
    jc = 2 * (1 + nc)
    m, J = nat.simulate(plan, prm, nc, nat.BMC_F64, m0, slots=slots, scales=scales, jc=jc)
    L = float(np.sum(gz * m.cpu().numpy()[..., jc]))
    if J is not None:
        L += float(np.sum(gj * J.cpu().numpy()))
    return L


def oracle_loss(records, ev, slots, scales, gz, gj):
    nc = len(records[0]["cest_k"])
    jc = 2 * (1 + nc)
    m0 = np.stack([O.initial_state(p) for p in records])
    m, dm = OC.simulate(pack_many(records), nc, ev, m0, slots=slots, rel_h=0.1)
    L = float(np.sum(gz * m[..., jc]))
    if slots:
        L += float(np.sum(gj * dm[..., jc] * np.asarray(scales)[None, None, :]))
    return L


def richardson(f, h):
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def shifted(ev, field, direction, eps):
    e = dict(ev)
    e[field] = ev[field] + eps * direction
    return e


@pytest.mark.parametrize("nc,P", [(3, 2), (2, 4)])
def test_more_pools_and_tangents_match_the_float64_oracle(nat, nc, P):
    """Three CEST pools and MT (16-lane teams), and four tangents (two launches whose gradients add)."""
    rng = np.random.default_rng(30 + nc + P)
    mk = {2: two_cest, 3: three_cest}[nc]
    recs = [mk(k=50.0, f=1.5e-3), mk(k=600.0, f=4e-3)]
    ev = train([-2.0, 3.5], n_pulses=2, n_samples=8)
    slots, scales = slots_for(nc, P), scales_for(nc, P)
    V, O_ = len(recs), ev["dt"].shape[0]
    gz, gj = rng.normal(size=(V, O_)), rng.normal(size=(V, O_, P))
    ga, gf = gpu_adjoint(nat, recs, ev, slots, scales, gz, gj)
    for field, g in (("rf_amp", ga), ("rf_freq", gf)):
        d = rng.normal(size=ev["dt"].shape)
        fd = richardson(lambda e: oracle_loss(recs, shifted(ev, field, d, e), slots, scales, gz, gj), 1e-2)
        ad = float(np.sum(g * d))
        assert abs(ad - fd) <= 2e-7 * np.sum(np.abs(g * d)) + 1e-12, (field, ad, fd)


@pytest.mark.parametrize("lineshape", ["superlorentzian", "lorentzian"])
def test_directional_derivative_matches_the_float64_oracle(nat, lineshape):
    rng = np.random.default_rng(3)
    recs = [three_pool(k=40.0, f=1.5e-3, lineshape=lineshape), three_pool(k=900.0, f=4e-3, t1w=1.6, b0i=0.1,
                                                                          lineshape=lineshape)]
    ev = train([-2.5, 1.0, 3.5])
    slots, scales = slots_for(1, 2), [100.0, 0.001]
    V, O_ = len(recs), ev["dt"].shape[0]
    gz, gj = rng.normal(size=(V, O_)), rng.normal(size=(V, O_, 2))
    ga, gf = gpu_adjoint(nat, recs, ev, slots, scales, gz, gj)
    for field, g, scale in (("rf_amp", ga, 1.0), ("rf_freq", gf, 1.0)):
        d = rng.normal(size=ev["dt"].shape) * scale
        fd = richardson(lambda e: oracle_loss(recs, shifted(ev, field, d, e), slots, scales, gz, gj), 1e-2)
        ad = float(np.sum(g * d))
        assert abs(ad - fd) <= 2e-7 * np.sum(np.abs(g * d)) + 1e-12, (field, ad, fd)


@pytest.mark.parametrize("nc,P", [(0, 1), (1, 0), (1, 1), (1, 2), (2, 2), (1, 3), (2, 6), (3, 2), (4, 4)])
def test_components_match_differences_of_the_kernel_forward(nat, nc, P):
    rng = np.random.default_rng(10 + 3 * nc + P)
    mk = {0: water_only, 1: three_pool, 2: two_cest, 3: three_cest, 4: four_cest}[nc]
    recs = [mk(), mk()] if nc == 0 else [mk(k=70.0, f=2e-3), mk(k=500.0, f=3e-3)]
    ev = train([-3.0, 3.5], n_pulses=2, n_samples=8)
    slots, scales = slots_for(nc, P), scales_for(nc, P)
    V, O_ = len(recs), ev["dt"].shape[0]
    gz, gj = rng.normal(size=(V, O_)), (rng.normal(size=(V, O_, P)) if P else None)
    ga, gf = gpu_adjoint(nat, recs, ev, slots, scales, gz, gj)
    T = ev["dt"].shape[1]
    picks = [(0, 0), (1, 1), (0, 5), (1, 9), (0, T - 3)]   # recovery, pulse samples, a gap
    for o, t in picks:
        for field, g in (("rf_amp", ga), ("rf_freq", gf)):
            d = np.zeros(ev["dt"].shape)
            d[o, t] = 1.0
            fd = richardson(lambda e: gpu_loss(nat, recs, shifted(ev, field, d, e), slots, scales, gz, gj), 1e-2)
            # relative 1e-6, plus the rounding of the differenced loss (|L| ~ 1, eps |L| / h ~ 1e-13)
            tol = 1e-6 * max(abs(fd), 1e-3 * float(np.max(np.abs(g)))) + 2e-12
            assert abs(g[o, t] - fd) <= tol, (field, o, t, g[o, t], fd)


def test_linear_in_the_cotangents(nat):
    rng = np.random.default_rng(4)
    recs = [three_pool(), three_pool(k=300.0)]
    ev = train([-2.0, 3.5, 5.0])
    slots, scales = slots_for(1, 2), [100.0, 0.001]
    V, O_ = 2, 3
    gz, gj = rng.normal(size=(V, O_)), rng.normal(size=(V, O_, 2))
    a1, f1 = gpu_adjoint(nat, recs, ev, slots, scales, gz, np.zeros_like(gj))
    a2, f2 = gpu_adjoint(nat, recs, ev, slots, scales, np.zeros_like(gz), gj)
    a3, f3 = gpu_adjoint(nat, recs, ev, slots, scales, gz, gj)
    np.testing.assert_allclose(a1 + a2, a3, rtol=1e-11, atol=1e-12 * np.abs(a3).max())
    np.testing.assert_allclose(f1 + f2, f3, rtol=1e-11, atol=1e-12 * np.abs(f3).max())


def test_events_before_a_reset_get_no_gradient(nat):
    """An ADC in the middle of the row reloads M0 (simulate.py:105-114): nothing before it reaches Z."""
    rng = np.random.default_rng(6)
    recs = [three_pool()]
    a = train([3.5, -1.0], n_pulses=1)
    b = train([3.5, -1.0], n_pulses=2)
    ev = {k: np.concatenate([a[k], b[k]], axis=1) for k in a}
    Ta = a["dt"].shape[1]
    slots, scales = slots_for(1, 2), [100.0, 0.001]
    gz, gj = rng.normal(size=(1, 2)), rng.normal(size=(1, 2, 2))
    ga, gf = gpu_adjoint(nat, recs, ev, slots, scales, gz, gj)
    assert np.all(ga[:, :Ta] == 0.0) and np.all(gf[:, :Ta] == 0.0)
    assert np.all(np.abs(ga[:, Ta + 1:Ta + 13]).max(axis=1) > 0)
    ga2, gf2 = gpu_adjoint(nat, recs, b, slots, scales, gz, gj)
    np.testing.assert_allclose(ga[:, Ta:], ga2, rtol=1e-12, atol=1e-14 * np.abs(ga2).max())


def test_float32_io_matches_float64(nat):
    rng = np.random.default_rng(8)
    recs = [three_pool(), three_pool(k=300.0, f=3e-3)]
    recs32 = [{k: (np.asarray(v, np.float32).astype(np.float64) if k != "mt_lineshape" else v) for k, v in p.items()}
              for p in recs]
    ev = train([-2.0, 3.5])
    slots, scales = slots_for(1, 2), [100.0, 0.001]
    gz, gj = rng.normal(size=(2, 2)), rng.normal(size=(2, 2, 2))
    a32, f32 = gpu_adjoint(nat, recs32, ev, slots, scales, gz, gj, dtype=np.float32)
    a64, f64 = gpu_adjoint(nat, recs32, ev, slots, scales, gz, gj, dtype=np.float64)
    assert np.max(np.abs(a32 - a64)) <= 1e-5 * np.abs(a64).max()
    assert np.max(np.abs(f32 - f64)) <= 1e-5 * np.abs(f64).max()


def test_unsupported_tangents_are_refused(nat):
    recs = [three_pool()]
    ev = train([3.5])
    with pytest.raises(nat.NativeError):
        gpu_adjoint(nat, recs, ev, [OC.slot_index("b0_inhom", None, 1)], [1.0], np.ones((1, 1)), np.ones((1, 1, 1)))
