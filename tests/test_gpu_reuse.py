"""The kept-pulse path of the simulate kernel (plan level 2: KEEP / REUSE ops, a pulse applied again as
R(d) P R(-d)) against the same kernel event by event (level 0), against the float64 C oracle (one Pade
exponential per event, phases as given), and against the golden fixture of two rows of a converted
scanner sequence that the reference ships. Through the C ABI, float64.
"""
import os

import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import four_pool_params, pack_many, rel_err

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def nat():
    from bmc_jax import _native
    _native.torch_cuda()
    return _native


def _ev_dict(ev):
    return {k: np.asarray(getattr(ev, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}


def run(nat, prm, m0, nc, ev, slots, level):
    import torch
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"], compress=level)
    m, J = nat.simulate(plan, nat.device_array(prm, torch.float64), nc, nat.BMC_F64,
                        nat.device_array(m0, torch.float64), slots=slots, scales=[1.0] * len(slots), jc=2 * (1 + nc))
    torch.cuda.synchronize()
    return m.cpu().numpy(), J.cpu().numpy(), plan


def _records(nc):
    if nc == 1:
        return [O.make_params(1.3, 0.08, pools=[dict(f=4e-3, k=k, dw=dw, t1=1.0, t2=0.015)], b0=2.8936,
                              b0_inhom=b0i, rel_b1=rb1)
                for k, dw, b0i, rb1 in ((900.0, 1.0, 0.0, 1.0), (120.0, 3.2, 0.05, 0.92), (2500.0, 1.9, -0.1, 1.07))]
    return [four_pool_params(k_am=k, f_am=2e-3, b0i=b0i, lineshape=ls)
            for k, b0i, ls in ((60.0, 0.0, "superlorentzian"), (400.0, 0.2, "lorentzian"))]


@pytest.mark.parametrize("nc", [1, 2])
def test_kept_pulses_match_the_event_by_event_scan_and_the_oracle(nat, nc):
    from bmc_jax.sequences import scanner_train
    ev = _ev_dict(scanner_train(np.array([-300.0, -4.0, -1.2, 0.3, 2.0, 3.5]), 1.5, n_pulses=5, n_samples=40,
                                lead_s=1.0, b0_T=2.8936))
    recs = _records(nc)
    prm = pack_many(recs)
    m0 = np.stack([O.initial_state(p, m_start=0.6) for p in recs])
    slots = [OC.slot_index("cest_k", 0, nc), OC.slot_index("cest_f", 0, nc), OC.slot_index("cest_r2", 0, nc),
             OC.slot_index("water_r1", None, nc), OC.slot_index("b0_inhom", None, nc)]
    m2, J2, plan2 = run(nat, prm, m0, nc, ev, slots, 2)
    m0_, J0, _ = run(nat, prm, m0, nc, ev, slots, 0)
    ops, _ = plan2.program()
    assert np.sum(ops[:, 0] == 3) == 4 * ev["type"].shape[0]      # four rotated reuses per row
    assert rel_err(m2, m0_) < 1e-12
    for p in range(len(slots)):
        assert rel_err(J2[..., p], J0[..., p]) < 1e-11, p
    mo, dmo = OC.simulate(prm, nc, ev, m0, slots=slots)
    assert rel_err(m2, mo) < 1e-12
    jz = 2 * (1 + nc)
    for p in range(len(slots)):
        assert rel_err(J2[..., p], dmo[..., p, jz]) < 1e-7, p


def test_scanner_rows_golden(nat):
    """Two rows of the reference's converted QUESP sequence (1 uT, 20 pulses at accumulated phases)."""
    g = np.load(os.path.join(GOLDEN, "scanner_rows.npz"))
    ev = {k: g[k] for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    slots = [int(s) for s in g["slots"]]
    m, J, plan = run(nat, g["params"], g["m0"], 1, ev, slots, 2)
    ops, _ = plan.program()
    assert np.sum(ops[:, 0] == 2) == 2 and np.sum(ops[:, 0] == 3) == 38
    assert rel_err(m, g["m"]) < 1e-12
    for p in range(len(slots)):
        assert rel_err(J[..., p], g["jac"][..., p]) < 1e-7, p


def test_the_lm_fit_on_a_kept_pulse_plan(nat):
    """bmc_lm_fit over a scanner train (persistent list sweep with the LDS keep slots) takes the same
    steps as over the event-by-event plan."""
    import torch
    from bmc_jax.sequences import scanner_train
    ev = _ev_dict(scanner_train(np.linspace(-4, 4, 9), 1.0, n_pulses=4, n_samples=30, lead_s=1.0))
    recs = _records(1)
    prm = pack_many(recs)
    m0 = np.stack([O.initial_state(p, m_start=0.557) for p in recs])
    slots = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1)]
    scales = [1000.0, 0.01]
    m, _, _ = run(nat, prm, m0, 1, ev, slots, 0)
    z = torch.as_tensor(m[..., 4], device="cuda")
    x0 = torch.as_tensor(np.array([[0.3, 0.8]] * len(recs)), device="cuda", dtype=torch.float64)
    out = []
    for level in (2, 0):
        plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"], compress=level)
        x, ct, _, nd = nat.lm_fit(plan, nat.device_array(prm, torch.float64), 1, nat.BMC_F64,
                                  nat.device_array(m0, torch.float64), slots, scales, [0.005, 0.001], [3.0, 3.0], 4, z,
                                  x0, 30, want_cost=True)
        out.append((x.cpu().numpy(), ct.cpu().numpy()))
    big = out[1][1] > 1e-12      # above the rounding floor the two plans take identical steps
    np.testing.assert_allclose(out[0][1][big], out[1][1][big], rtol=1e-6)
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=1e-7)
    truth = np.array([[p["cest_k"][0] / 1000.0, p["cest_f"][0] / 0.01] for p in recs])
    np.testing.assert_allclose(out[0][0], truth, rtol=1e-3)
