"""bmc_lm_fit_noalloc (the allocation-free LM entry point) against bmc_lm_fit through the C ABI.

Both run the same kernels, so the fitted values, cost trajectories and iteration counts must be equal
bit for bit. Two calls in a row, the second smaller, show that the fixed workspace keeps nothing a call
depends on; a call too large for the workspace and a graph request are refused with a message."""
import ctypes

import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC

pytestmark = pytest.mark.gpu


def _problem(n_vox, ppm, b1s, seed, pulsed=False):
    """Seeded voxels of the configuration-1 pool model, their spectra (float64 oracle) and a start."""
    from bmc_jax import _native as nat
    import torch
    rng = np.random.default_rng(seed)
    recs = [O.make_params(1.3, 0.075, pools=[dict(f=float(rng.uniform(3e-4, 3e-3)),
                                                   k=float(np.exp(rng.uniform(np.log(20), np.log(800)))),
                                                   dw=3.5, t1=1.3, t2=0.1)], b0=3.0)
            for _ in range(n_vox)]
    if pulsed:
        ev = O.pulsed_train(ppm, b1_peak_uT=2.0, n_pulses=3, n_samples=12, recovery_s=1.0, b0_T=3.0,
                            dtype=np.float64)
    else:
        parts = [O.cw_saturation(ppm, b1_uT=b, tsat_s=2.0, b0_T=3.0) for b in b1s]
        ev = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    prm = np.stack([OC.pack(p) for p in recs])
    m0 = np.stack([O.initial_state(p) for p in recs])
    slots = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1), OC.slot_index("water_r2", None, 1)]
    m, _ = OC.simulate(prm, 1, ev, m0)
    z = m[:, :, 4]
    plan = nat.Plan(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"])
    dev = lambda a: nat.device_array(a, torch.float64)
    x0 = np.stack([np.array([1.5, 0.6, 1.2]) * rng.uniform(0.8, 1.2, 3) for _ in range(n_vox)])
    return dict(plan=plan, prm=dev(prm), m0=dev(m0), z=dev(z), x0=torch.as_tensor(x0), slots=slots,
                scales=[100.0, 0.001, 10.0], lo=[0.05, 0.01, 0.1], hi=[20.0, 30.0, 10.0])


def _fit(q, noalloc, flags=0, n_iter=8):
    import torch
    from bmc_jax import _native as nat
    x, ct, _, nd = nat.lm_fit(q["plan"], q["prm"], 1, nat.BMC_F64, q["m0"], q["slots"], q["scales"], q["lo"],
                              q["hi"], 4, q["z"], q["x0"], n_iter, flags=flags, early_ftol=1e-12, early_rms=1e-9,
                              want_cost=True, noalloc=noalloc)
    torch.cuda.synchronize()
    return x.cpu().numpy(), ct.cpu().numpy(), nd.cpu().numpy()


@pytest.fixture(scope="module")
def problems():
    from bmc_jax import _native as nat
    nat.torch_cuda()
    return dict(
        big=_problem(96, np.linspace(-6, 6, 61), (0.5, 1.0, 2.0, 4.0), 1),
        small=_problem(17, np.linspace(-5, 5, 31), (1.0, 3.0), 2),
        train=_problem(9, np.linspace(-4, 4, 9), (), 3, pulsed=True),
    )


@pytest.mark.parametrize("flags", [0, 1 | 4])
def test_two_calls_in_a_row_match_the_allocating_entry_point(problems, flags):
    """Large then small (CW ladders: the reverse-accumulation Jacobian), then a pulsed train (the
    forward-mode kernel); with and without early exit and the start check."""
    for name in ("big", "small", "train"):
        q = problems[name]
        xa, ca, na = _fit(q, noalloc=True, flags=flags)
        xb, cb, nb = _fit(q, noalloc=False, flags=flags)
        np.testing.assert_array_equal(xa, xb, err_msg=name)
        np.testing.assert_array_equal(ca, cb, err_msg=name)
        np.testing.assert_array_equal(na, nb, err_msg=name)
        assert np.all(np.isfinite(xa)), name
        # the cost only falls; once converged, the Jacobian sweep's own evaluation of the same point (another
        # Taylor / squaring split of exp) may sit a rounding above the trial's
        assert np.all(ca[:, -1] <= ca[:, 0] * (1 + 1e-6) + 1e-24), (name, ca[:, [0, -1]])


def test_refusals_come_back_as_errors(problems):
    import torch
    from bmc_jax import _native as nat
    q = problems["small"]
    with pytest.raises(nat.NativeError, match="BMC_LM_GRAPH"):
        _fit(q, noalloc=True, flags=nat.LM_GRAPH)
    # 16384 voxels x 62 rows x 16 values need ~146 MB: above the fixed 64 MiB, refused before any launch
    # (the placeholder buffers are never touched)
    V, P = 16384, 16
    L = nat.lib()
    arr = lambda ct, v: (ct * len(v))(*v)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    ptr = ctypes.c_void_p(dummy.data_ptr())
    rc = L.bmc_lm_fit_noalloc(q["plan"].handle, V, 1, nat.BMC_F64, ptr, ptr, P, arr(ctypes.c_int32, [0] * P),
                              arr(ctypes.c_double, [1.0] * P), arr(ctypes.c_double, [0.0] * P),
                              arr(ctypes.c_double, [1.0] * P), 4, ptr, ptr, 5, 1e-2, None, 0, 0.0, 0.0, None, None,
                              None, nat.stream_ptr())
    assert rc == -1 and b"workspace" in L.bmc_last_error()
