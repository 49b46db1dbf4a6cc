"""Parity of the HIP forward kernel (bmc_simulate, through the C ABI) with the float64 C oracle.

The oracle takes one float64 Pade-13 exponential per event with no reuse and derives the generator by
central differences; the kernel uses Taylor columns, squarings, chained repeated pulses and analytic
generator derivatives. Agreement at the 1e-11 level (state) and 1e-7 (Jacobian, limited by the
oracle's differenced generator) therefore checks the whole path, not a shared formula.
"""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import (as_dtype, cfg1_events, cfg1_params, four_pool_params, pack_many, rel_err, small_train)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nat():
    from bmc_jax import _native
    _native.torch_cuda()
    return _native


def run_gpu(nat, records, ev, slots, dtype=np.float64, m0=None, dm0=None, jc=None):
    import torch
    nc = len(records[0]["cest_k"])
    N = 3 * (1 + nc) + 1
    code = nat.BMC_F64 if dtype == np.float64 else nat.BMC_F32
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    prm = nat.device_array(pack_many(records).astype(dtype), tdt)
    if m0 is None:
        m0 = np.stack([O.initial_state(p) for p in records])
    m0d = nat.device_array(np.asarray(m0, dtype), tdt)
    dm0d = None if dm0 is None else nat.device_array(np.asarray(dm0, dtype), tdt)
    jc = 2 * (1 + nc) if jc is None else jc
    m, J = nat.simulate(plan, prm, nc, code, m0d, dm0=dm0d, slots=slots, scales=[1.0] * len(slots), jc=jc)
    torch.cuda.synchronize()
    return m.double().cpu().numpy(), (None if J is None else J.double().cpu().numpy()), plan


def run_oracle(records, ev, slots, m0=None, dm0=None):
    nc = len(records[0]["cest_k"])
    if m0 is None:
        m0 = np.stack([O.initial_state(p) for p in records])
    m, dm = OC.simulate(pack_many(records), nc, ev, m0, slots=slots, dm0=dm0)
    return m, dm


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_cw_configuration_1(nat, dtype):
    """Configuration 1 (2-pool CW, 61 offsets): Z and the (k, f) Jacobian."""
    p = cfg1_params()
    ev = cfg1_events()
    nc = 1
    slots = [OC.slot_index("cest_k", 0, nc), OC.slot_index("cest_f", 0, nc)]
    rec = p
    if dtype == np.float32:   # the float32 problem: every input rounded to float32, then exact oracle
        rec = {k: (np.asarray(v, np.float32).astype(np.float64) if k != "mt_lineshape" else v) for k, v in p.items()}
    m_g, J_g, _ = run_gpu(nat, [rec], as_dtype(ev, dtype), slots, dtype=dtype)
    m_o, dm_o = run_oracle([rec], as_dtype(ev, dtype).copy(), slots)
    mz = 2 * (1 + nc)
    z_g, z_o = m_g[0, :, mz], m_o[0, :, mz]
    big = np.abs(z_o) > 0.05
    tol_z = 1e-12 if dtype == np.float64 else 2e-7
    assert np.max(np.abs(z_g - z_o)[big] / np.abs(z_o)[big]) < tol_z
    assert np.max(np.abs(z_g - z_o)) < (1e-12 if dtype == np.float64 else 1e-7)
    J_o = dm_o[0, :, :, mz]
    tol_j = 1e-7 if dtype == np.float64 else 2e-6
    for c in range(2):
        assert rel_err(J_g[0, :, c], J_o[:, c]) < tol_j, f"column {c}"
    # the whole state, not only water Mz
    assert np.max(np.abs(m_g - m_o)) < (1e-12 if dtype == np.float64 else 2e-7)


ALL_FIELDS = [("water_r1", None), ("water_r2", None), ("water_f", None), ("mt_k", None), ("mt_f", None),
              ("mt_dw", None), ("mt_r1", None), ("mt_r2", None), ("b0", None), ("gamma", None),
              ("b0_inhom", None), ("rel_b1", None), ("cest_k", 0), ("cest_f", 0), ("cest_dw", 0), ("cest_r1", 0),
              ("cest_r2", 0), ("cest_k", 1), ("cest_f", 1), ("cest_dw", 1), ("cest_r1", 1), ("cest_r2", 1)]


@pytest.mark.parametrize("lineshape", ["superlorentzian", "lorentzian"])
def test_pulsed_train_every_tangent(nat, lineshape):
    """4 pools with MT, a small pulsed train (recovery, pulses, gaps, spoiler, ADC), offsets on both
    MT lineshape branches; the derivative with respect to every packed field."""
    ppm = np.array([-6.0, -2.7, -2.2, 0.3, 2.0, 3.5])   # -2.7 / -2.2 sit on the cubic-Hermite bridge
    ev = small_train(ppm)
    recs = [four_pool_params(k_am=50.0, b0i=0.05, lineshape=lineshape),
            four_pool_params(k_am=1200.0, f_am=5e-3, k_amine=3000.0, t1w=2.2, t2w=0.1, b0i=-0.2,
                             lineshape=lineshape, rel_b1=0.9)]
    nc = 2
    slots = [OC.slot_index(f, i, nc) for f, i in ALL_FIELDS]
    mz = 2 * (1 + nc)
    for part in (slots[:12], slots[12:]):
        m_g, J_g, plan = run_gpu(nat, recs, ev, part)
        m_o, dm_o = run_oracle(recs, ev, part)
        assert np.max(np.abs(m_g - m_o)) < 1e-11
        J_o = dm_o[..., mz]
        for c in range(len(part)):
            scale = max(np.max(np.abs(J_o[:, :, c])), 1e-12)
            err = np.max(np.abs(J_g[:, :, c] - J_o[:, :, c])) / scale
            assert err < 3e-6 or np.max(np.abs(J_g[:, :, c] - J_o[:, :, c])) < 1e-10, \
                f"slot {part[c]}: rel {err:.2e}"
    st = plan.stats()
    assert st["n_built"] < ev["dt"].size / 2, "the repeated pulses were not chained"


def test_adc_reset_and_spoiler_in_the_middle(nat):
    """An ADC halfway resets the state; a spoiler mid-row zeroes transverse magnetisation."""
    ppm = np.array([-4.0, 1.0, 3.5])
    a = small_train(ppm, n_pulses=2, n_samples=8)
    b = small_train(ppm, n_pulses=3, n_samples=8)
    ev = {k: np.concatenate([a[k], b[k]], axis=1) for k in a}
    recs = [four_pool_params()]
    nc = 2
    slots = [OC.slot_index("cest_k", 0, nc), OC.slot_index("b0_inhom", None, nc)]
    m_g, J_g, _ = run_gpu(nat, recs, ev, slots)
    m_o, dm_o = run_oracle(recs, ev, slots)
    assert np.max(np.abs(m_g - m_o)) < 1e-11
    assert rel_err(J_g[0], dm_o[0, :, :, 6]) < 1e-6
    # the second half alone from M0 gives the same answer (the reset is exact)
    m_b, _, _ = run_gpu(nat, recs, b, slots)
    assert np.max(np.abs(m_b - m_g)) < 1e-12


def test_front_padding_is_a_no_op(nat):
    """Zero-duration DELAYs before the sequence change nothing, to the last bit (test_experiment.py:29-36)."""
    ev = cfg1_events()
    pad = {k: np.concatenate([np.zeros((61, 5), v.dtype) if k != "adc_idx" else -np.ones((61, 5), v.dtype), v],
                             axis=1) for k, v in ev.items()}
    recs = [cfg1_params()]
    m1, _, _ = run_gpu(nat, recs, ev, [])
    m2, _, _ = run_gpu(nat, recs, pad, [])
    assert np.array_equal(m1, m2)


def test_tail_padding_returns_the_start(nat):
    """Appending no-ops after the ADC returns M0 (test_experiment.py:39-69): last_M semantics."""
    ev = cfg1_events()
    tail = {k: np.concatenate([v, np.zeros((61, 4), v.dtype) if k != "adc_idx" else -np.ones((61, 4), v.dtype)],
                              axis=1) for k, v in ev.items()}
    p = cfg1_params()
    m, _, _ = run_gpu(nat, [p], tail, [])
    assert np.max(np.abs(m[0] - O.initial_state(p)[None, :])) < 1e-15


def test_nan_rule_and_recovery_after_reset(nat):
    """An event needing more than 16 reference squarings gives NaN (expm.py:83); an ADC reset clears it."""
    p = cfg1_params()
    ev = O.cw_saturation([-300.0, 3.5], b1_uT=1.0, tsat_s=5.0, b0_T=3.0)
    m, _, _ = run_gpu(nat, [p], ev, [], dtype=np.float32)
    assert np.isnan(m[0, 0]).all() and np.isfinite(m[0, 1]).all()
    two = {k: np.concatenate([v, v[::-1]], axis=1) for k, v in ev.items()}   # row 0: NaN block, ADC, then fine
    m2, _, _ = run_gpu(nat, [p], two, [], dtype=np.float32)
    assert np.isfinite(m2[0, 0]).all()


def test_initial_state_tangent_seed(nat):
    """dM0 seeds propagate: d Z / d m_start with M0 = m_start * M_eq (joint.py:137)."""
    p = cfg1_params()
    ev = cfg1_events()
    m_eq = O.initial_state(p)
    ms = 0.6
    m_g, J_g, _ = run_gpu(nat, [p], ev, [-1], m0=(m_eq * ms)[None], dm0=m_eq[None, None, :])
    m_o, dm_o = run_oracle([p], ev, [], m0=(m_eq * ms)[None])
    h = 1e-6
    m_p, _ = run_oracle([p], ev, [], m0=(m_eq * (ms + h))[None])
    m_m, _ = run_oracle([p], ev, [], m0=(m_eq * (ms - h))[None])
    fd = (m_p - m_m)[0, :, 4] / (2 * h)
    assert rel_err(J_g[0, :, 0], fd) < 1e-7


def test_configuration_2_full_train_golden(nat):
    """The full pulsed train (50 x 200 Gaussian steps, 10,053 events) on the golden slice: 3 voxels x 6
    offsets, 5 free values, against the committed float64 oracle fixture (tests/golden/make_golden.py)."""
    import os
    import torch
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg2_slice.npz"))
    ev = O.pulsed_train(g["ppm"], b1_peak_uT=2.0, n_pulses=50, pulse_s=0.02, n_samples=200, gap_s=0.02,
                        recovery_s=3.0, b0_T=3.0, dtype=np.float32)
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    st = plan.stats()
    assert st["n_built"] == 6 * 204, st         # recovery + 201-event block + spoiler + ADC per offset
    prm = nat.device_array(g["params"], torch.float64)
    m0 = nat.device_array(g["m0"], torch.float64)
    slots = [int(s) for s in g["slots"]]
    m, J = nat.simulate(plan, prm, 2, nat.BMC_F64, m0, slots=slots, scales=[1.0] * 5, jc=6)
    z = m.cpu().numpy()[..., 6]
    J = J.cpu().numpy()
    assert np.max(np.abs(z - g["z"])) < 1e-11
    for c in range(5):
        ref = g["jac"][..., c]
        assert np.max(np.abs(J[..., c] - ref)) < 1e-6 * np.abs(ref).max(), f"column {c}"


def test_configuration_2_split_launch_golden(nat):
    """The benchmark's launch shape: past 2^17 (voxel, offset) teams the five tangents run as passes of two
    and then a launch of their own for the fifth (csrc/bmc_sim.hip launch_sim), with the palindromic pulses
    factorised. The golden slice's 3 voxels tiled to 21,846 voxels x 6 offsets (131,076 teams), every copy
    against the committed float64 oracle fixture, in float64 and with the benchmark's float32 I/O."""
    import os
    import torch
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg2_slice.npz"))
    ev = O.pulsed_train(g["ppm"], b1_peak_uT=2.0, n_pulses=50, pulse_s=0.02, n_samples=200, gap_s=0.02,
                        recovery_s=3.0, b0_T=3.0, dtype=np.float32)
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    reps = 7282
    n_vox = 3 * reps
    assert n_vox * len(g["ppm"]) >= 1 << 17
    slots = [int(s) for s in g["slots"]]
    for dtype, tz, tj in ((nat.BMC_F64, 1e-11, 1e-6), (nat.BMC_F32, 2e-6, 2e-4)):
        tdt = torch.float64 if dtype == nat.BMC_F64 else torch.float32
        prm = nat.device_array(np.tile(g["params"], (reps, 1)), tdt)
        m0 = nat.device_array(np.tile(g["m0"], (reps, 1)), tdt)
        m, J = nat.simulate(plan, prm, 2, dtype, m0, slots=slots, scales=[1.0] * 5, jc=6)
        z = m.cpu().numpy().astype(np.float64)[..., 6].reshape(reps, 3, -1)
        J = J.cpu().numpy().astype(np.float64).reshape(reps, 3, len(g["ppm"]), 5)
        assert np.max(np.abs(z - g["z"][None])) < tz
        for c in range(5):
            ref = g["jac"][..., c]
            assert np.max(np.abs(J[..., c] - ref[None])) < tj * np.abs(ref).max(), (dtype, c)
        del m, J, prm, m0


@pytest.mark.parametrize("n_tan", [0, 1])
@pytest.mark.parametrize("level", [2, 3])
def test_two_pool_kernels_with_staged_squarings(nat, n_tan, level):
    """Kernels of 12-lane teams (two CEST pools) with at most one tangent per pass run the squarings'
    products through the team's LDS stage (staged_event in csrc/bmc_device.h): the recovery delay as a
    step of its own, the gaps inside the chained pulses, and CW rows (one event, about ten squarings) with
    forward-mode tangents. Against the float64 oracle, both plan levels."""
    import torch
    recs = [four_pool_params(k_am=300.0, f_am=2e-3, b0i=0.05),
            four_pool_params(k_am=2500.0, f_am=6e-3, t1w=2.1, b0i=-0.12, rel_b1=0.9)]
    nc = 2
    slots = [OC.slot_index("cest_f", 0, nc)] if n_tan else []
    prm = nat.device_array(pack_many(recs), torch.float64)
    m0 = nat.device_array(np.stack([O.initial_state(p) for p in recs]), torch.float64)
    for ev, fwd in ((small_train(np.linspace(-5.0, 5.0, 7)), False),
                    (O.cw_saturation(np.linspace(-6.0, 6.0, 9), b1_uT=1.5, tsat_s=2.0, b0_T=3.0), True)):
        plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"], compress=level)
        m, J = nat.simulate(plan, prm, nc, nat.BMC_F64, m0, slots=slots, scales=[1.0] * len(slots),
                            jc=2 * (1 + nc), forward_jacobian=fwd)
        torch.cuda.synchronize()
        mo, dmo = run_oracle(recs, ev, slots)
        assert np.max(np.abs(m.cpu().numpy() - mo)) < 1e-11
        if n_tan:
            Jg, Jo = J.cpu().numpy()[..., 0], dmo[..., 0, 2 * (1 + nc)]
            assert np.max(np.abs(Jg - Jo)) < 1e-7 * np.abs(Jo).max()
