"""bmc_lm_fit's start check (BMC_LM_START_CHECK with early exit): problems whose start already meets the
rms tolerance stop before their first Jacobian and keep their start; the others take exactly the steps
of a run without the check. Without the check, early exit stops a converged start on its first Jacobian
sweep's state, before any step. Through the C ABI, float64, configuration 1's CW spectrum."""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import cfg1_events, cfg1_params, pack_many

pytestmark = pytest.mark.gpu


def test_start_check_stops_converged_starts_only():
    import torch
    from bmc_jax import _native as nat
    nat.torch_cuda()
    p = cfg1_params()
    ev = cfg1_events()
    V = 6
    recs = [p] * V
    slots = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1)]
    scales = [100.0, 0.001]
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    prm = nat.device_array(pack_many(recs), torch.float64)
    m0 = nat.device_array(np.stack([O.initial_state(q) for q in recs]), torch.float64)
    m, _ = nat.simulate(plan, prm, 1, nat.BMC_F64, m0, want_jac=False)
    z = m[..., 4].contiguous()
    truth = np.array([50.0 / 100.0, 9.01e-4 / 0.001])
    x0 = np.tile(truth, (V, 1))
    x0[1::2] *= np.array([1.6, 0.5])            # odd problems start away from the truth
    x0_t = torch.as_tensor(x0, device="cuda")
    lo, hi = [0.05, 0.01], [20.0, 20.0]
    runs = {}
    for name, flags in (("check", nat.LM_EARLY_EXIT | nat.LM_START_CHECK), ("plain", nat.LM_EARLY_EXIT)):
        x, ct, _, nd = nat.lm_fit(plan, prm, 1, nat.BMC_F64, m0, slots, scales, lo, hi, 4, z, x0_t, 20, flags=flags,
                                  early_ftol=0.0, early_rms=1e-9, want_cost=True)
        runs[name] = (x.cpu().numpy(), ct.cpu().numpy(), nd.cpu().numpy())
    x, ct, nd = runs["check"]
    assert np.all(nd[0::2] == 0) and np.all(nd[1::2] > 0)
    np.testing.assert_array_equal(x[0::2], x0[0::2])
    assert np.all(ct[0::2] < 1e-24)
    xp, ctp, ndp = runs["plain"]
    # without the separate sweep the first Jacobian sweep's own state stops a converged start: no step
    assert np.all(ndp[0::2] == 0)
    np.testing.assert_array_equal(xp[0::2], x0[0::2])
    np.testing.assert_array_equal(nd[1::2], ndp[1::2])
    np.testing.assert_array_equal(x[1::2], xp[1::2])
    np.testing.assert_allclose(x[1::2], np.tile(truth, (V // 2, 1)), rtol=1e-5)   # rms 1e-9 leaves ~1e-6
