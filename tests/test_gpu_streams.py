"""Stream ordering of the graph-captured LM on the caller's null stream (torch's default stream).

The null stream cannot be captured, so bmc_lm_fit and bmc_ubatch_lm_fit run a graph-captured iteration on
a stream of their own. That stream has to wait for the work the caller already queued on the null stream:
here the start values come out of a long null-stream reduction launched just before the call. The graph
run must give exactly what the uncaptured run (on the null stream itself) gives."""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC
from helpers import cfg1_events, cfg1_params, pack_many

pytestmark = pytest.mark.gpu


def test_graph_lm_waits_for_queued_null_stream_work():
    import torch
    from bmc_jax import _native as nat
    nat.torch_cuda()
    assert torch.cuda.current_stream().cuda_stream == 0      # the caller's stream is the null stream
    p = cfg1_params()
    ev = cfg1_events()
    V = 512
    recs = [p] * V
    slots = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1)]
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    prm = nat.device_array(pack_many(recs), torch.float64)
    m0 = nat.device_array(np.stack([O.initial_state(q) for q in recs]), torch.float64)
    m, _ = nat.simulate(plan, prm, 1, nat.BMC_F64, m0, want_jac=False)
    z = m[..., 4].contiguous()
    start = torch.as_tensor(np.tile([1.3, 0.6], (V, 1)), device="cuda")
    big = torch.ones(1 << 27, device="cuda", dtype=torch.float64)
    out = {}
    for graph in (False, True):
        torch.cuda.synchronize()
        # x0 is ready only after a long reduction queued on the null stream right before the call
        x0 = start + 0.0 * big.sum()
        x, ct, _, nd = nat.lm_fit(plan, prm, 1, nat.BMC_F64, m0, slots, [100.0, 0.001], [0.05, 0.01], [20.0, 20.0],
                                  4, z, x0, 12, flags=nat.LM_GRAPH if graph else 0, want_cost=True)
        torch.cuda.synchronize()
        out[graph] = (x.cpu().numpy(), ct.cpu().numpy(), nd.cpu().numpy())
    for a, b in zip(out[False], out[True]):
        np.testing.assert_array_equal(a, b)
    assert np.all(out[True][2] == 12)
