"""The palindrome factorisation of chained pulses (plan level 3; mirror_chain in csrc/bmc_device.h).

A symmetric pulse shape at one RF phase makes a chained run read the same backwards, and its propagator
is then Q^R Q with Q^R = S Q^T S^-1 built from the first half. These tests compare the factorised kernel
with the float64 C oracle (one Pade-13 exponential per event, no reuse) and with the same plan compiled
at level 2 (every event chained), for every tangent kind, pool counts 0 and 2-4, RF phase 0 and not 0,
decoupled pools, and the cases that must fall back to event-by-event chaining (a zero fraction with a
nonzero rate, a half whose events need squarings), which then agree with level 2 bit for bit.
"""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import pack_many

pytestmark = pytest.mark.gpu

FIELDS = [("water_r1", None), ("water_r2", None), ("water_f", None), ("mt_k", None), ("mt_f", None),
          ("mt_dw", None), ("mt_r1", None), ("mt_r2", None), ("b0", None), ("gamma", None), ("b0_inhom", None),
          ("rel_b1", None)]
POOL_FIELDS = ("cest_k", "cest_f", "cest_dw", "cest_r1", "cest_r2")


@pytest.fixture(scope="module")
def nat():
    from bmc_jax import _native
    _native.torch_cuda()
    return _native


def train(ppm, phase=0.0, n_samples=64):
    """Three Gaussian pulses (symmetric samples) with gaps, a spoiler and the readout; RF at `phase`."""
    ev = O.pulsed_train(np.asarray(ppm, float), b1_peak_uT=2.0, n_pulses=3, pulse_s=0.02, n_samples=n_samples,
                        gap_s=0.02, recovery_s=1.0, b0_T=3.0, dtype=np.float64)
    ev = dict(ev)
    ev["rf_phase"] = np.where(ev["type"] == 1, phase, 0.0)
    return ev


def pools(nc, f0=1e-3):
    return [dict(f=f0 * (1 + 2 * i), k=60.0 * (4 ** i), dw=3.5 - 0.8 * i, t1=1.3, t2=0.1 - 0.015 * i)
            for i in range(nc)]


def run(nat, recs, ev, slots, level):
    import torch
    nc = len(recs[0]["cest_k"])
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"], compress=level)
    prm = nat.device_array(pack_many(recs), torch.float64)
    m0 = nat.device_array(np.stack([O.initial_state(p) for p in recs]), torch.float64)
    m, J = nat.simulate(plan, prm, nc, nat.BMC_F64, m0, slots=slots, scales=[1.0] * len(slots), jc=2 * (1 + nc))
    torch.cuda.synchronize()
    return m.cpu().numpy(), (None if J is None else J.cpu().numpy()), plan


def col_err(J, ref):
    """Largest error of each Jacobian column relative to that column's largest oracle entry."""
    return [float(np.max(np.abs(J[..., c] - ref[..., c])) / max(np.max(np.abs(ref[..., c])), 1e-12))
            for c in range(ref.shape[-1])]


@pytest.mark.parametrize("phase", [0.0, 0.7])
@pytest.mark.parametrize("nc", [0, 2, 3, 4])
def test_factorised_pulses_match_the_oracle(nat, nc, phase):
    ppm = [-6.0, -2.6, 0.3, 2.0, 3.5]
    ev = train(ppm, phase)
    recs = [O.make_params(1.3, 0.07, pools=pools(nc), b0=3.0, b0_inhom=0.05,
                          mt=dict(f=0.1, k=30.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape="superlorentzian")),
            O.make_params(2.1, 0.1, pools=pools(nc, 4e-3), b0=3.0, b0_inhom=-0.2, rel_b1=0.9,
                          mt=dict(f=0.05, k=20.0, dw=-2.0, t1=1.0, t2=2e-5, lineshape="lorentzian"))]
    names = FIELDS + [(f, i) for i in range(nc) for f in POOL_FIELDS]
    slots = [OC.slot_index(f, i, nc) for f, i in names]
    mz = 2 * (1 + nc)
    for lo in range(0, len(slots), 10):
        part = slots[lo:lo + 10]
        m3, J3, plan = run(nat, recs, ev, part, 3)
        m2, J2, _ = run(nat, recs, ev, part, 2)
        mo, dmo = OC.simulate(pack_many(recs), nc, ev, np.stack([O.initial_state(p) for p in recs]), slots=part)
        ops, _ = plan.program()
        assert np.any(ops[:, 4] == 32), "the pulses were not marked as palindromes"
        assert np.max(np.abs(m3 - mo)) < 1e-11
        assert np.max(np.abs(m3 - m2)) < 1e-12
        assert np.any(m3 != m2), "the factorised path did not run"
        e_o = col_err(J3, dmo[..., mz])
        e_2 = col_err(J3, J2)
        for c, name in enumerate(names[lo:lo + 10]):
            assert e_o[c] < 3e-6 or np.max(np.abs(J3[..., c] - dmo[..., c, mz])) < 1e-10, \
                f"{name}: {e_o[c]:.2e} against the oracle"
            assert e_2[c] < 1e-9, f"{name}: {e_2[c]:.2e} against level 2"


def test_decoupled_pools_and_fallbacks(nat):
    """One launch, four voxels of two pools: (a) MT absent and the second pool decoupled (f = k = 0), which
    factorise; (b) the second pool with f = 0 but k != 0, and (c) a pool rate so high that the half's
    events need squarings, which fall back to event-by-event chaining and match level 2 bit for bit;
    (d) a plain voxel beside them. Tangents include the decoupled pool's fraction and R1 and MT's fraction
    with MT absent."""
    nc = 2
    ev = train([-4.0, -1.0, 0.5, 3.5], phase=0.3)
    base = pools(2)
    dec = [base[0], dict(f=0.0, k=0.0, dw=2.0, t1=1.3, t2=0.05)]
    bad = [base[0], dict(f=0.0, k=40.0, dw=2.0, t1=1.3, t2=0.05)]
    hot = [dict(base[0], k=2.5e4), base[1]]
    recs = [O.make_params(1.3, 0.07, pools=dec, b0=3.0, n_cest=2),
            O.make_params(1.3, 0.07, pools=bad, b0=3.0, n_cest=2),
            O.make_params(1.3, 0.07, pools=hot, b0=3.0, n_cest=2),
            O.make_params(1.5, 0.08, pools=base, b0=3.0, b0_inhom=0.1, n_cest=2,
                          mt=dict(f=0.1, k=30.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape="superlorentzian"))]
    names = [("cest_k", 0), ("cest_f", 1), ("cest_f", 0), ("water_r1", None), ("mt_f", None), ("cest_r1", 1)]
    slots = [OC.slot_index(f, i, nc) for f, i in names]
    m3, J3, _ = run(nat, recs, ev, slots, 3)
    m2, J2, _ = run(nat, recs, ev, slots, 2)
    mo, dmo = OC.simulate(pack_many(recs), nc, ev, np.stack([O.initial_state(p) for p in recs]), slots=slots)
    assert np.max(np.abs(m3 - mo)) < 1e-11
    e_o = col_err(J3, dmo[..., 2 * (1 + nc)])
    assert max(e_o) < 3e-6, e_o
    assert np.any(m3[0] != m2[0]) and np.any(m3[3] != m2[3]), "eligible voxels were not factorised"
    for v in (1, 2):
        np.testing.assert_array_equal(m3[v], m2[v])
        np.testing.assert_array_equal(J3[v], J2[v])
