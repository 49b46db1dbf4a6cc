"""Host side of every device call: plan cache, packing, device buffers, launches.

The kernels (libbmc_hip.so) do the numerics; this module turns the reference's NumPy-shaped arguments
(SimParams, SimEvent, M_init) into packed device buffers and brings the results back.
"""
from __future__ import annotations

import hashlib
from collections import OrderedDict

import numpy as np

from . import _native, settings
from .simulation.params import SimParams, batch_size, n_cest_of, pack

_PLANS: "OrderedDict[tuple, _native.Plan]" = OrderedDict()
_PLAN_CACHE_SIZE = 32


def _events_key(events, reset_after_adc, compress):
    h = hashlib.blake2b(digest_size=16)
    for f in ("type", "dt", "rf_amp", "rf_phase", "rf_freq"):
        a = np.ascontiguousarray(np.asarray(getattr(events, f)))
        h.update(f.encode())
        h.update(str(a.shape).encode())
        h.update(a.astype(np.float64 if f != "type" else np.int32).tobytes())
    return (h.hexdigest(), bool(reset_after_adc), _native._level(compress))


def plan_for(events, reset_after_adc=True, compress=True) -> _native.Plan:
    """A device plan for an event table, cached on its exact contents. compress: True / 3 (runs,
    phase-shifted pulses reused and palindromic runs factorised), 2 (no factorisation), 1 (runs only),
    False / 0 (every event its own step)."""
    key = _events_key(events, reset_after_adc, compress)
    plan = _PLANS.get(key)
    if plan is not None:
        _PLANS.move_to_end(key)
        return plan
    t = np.asarray(events.type)
    if t.ndim != 2:
        raise ValueError(f"events must be batched (n_offsets, n_events); got shape {t.shape}")
    plan = _native.Plan(events.type, events.dt, events.rf_amp, events.rf_phase, events.rf_freq,
                        reset_after_adc=reset_after_adc, compress=compress)
    _PLANS[key] = plan
    while len(_PLANS) > _PLAN_CACHE_SIZE:
        _PLANS.popitem(last=False)
    return plan


def dtype_code(dtype=None):
    dtype = np.dtype(dtype or settings.float_dtype())
    if dtype == np.float64:
        return _native.BMC_F64, np.float64
    if dtype == np.float32:
        return _native.BMC_F32, np.float32
    raise ValueError(f"unsupported dtype {dtype}")


def _torch_dtype(np_dtype):
    torch = _native.torch_cuda()
    return torch.float64 if np_dtype == np.float64 else torch.float32


def check_pools(nc):
    if nc > _native.MAX_CEST:
        raise NotImplementedError(f"{nc} CEST pools: the kernels are compiled for at most {_native.MAX_CEST}")


def prepare(params: SimParams, M_init, n_voxels=None, dtype=None):
    """Packed params and M_init as device tensors. Returns (prm, m0, nc, V, code, np_dtype)."""
    code, npd = dtype_code(dtype)
    nc = n_cest_of(params)
    check_pools(nc)
    V = batch_size(params)
    m0 = np.asarray(M_init)
    if V is None:
        V = n_voxels if n_voxels is not None else (m0.shape[0] if m0.ndim == 2 else 1)
    N = 3 * (1 + nc) + 1
    if m0.shape[-1] != N:
        raise ValueError(f"M_init has length {m0.shape[-1]}, the parameters need {N}")
    m0 = np.broadcast_to(m0, (V, N))
    prm = pack(params, V, dtype=npd)
    tdt = _torch_dtype(npd)
    return (_native.device_array(prm, tdt), _native.device_array(m0.astype(npd), tdt), nc, V, code, npd)


def simulate_states(params: SimParams, events, M_init, n_voxels=None, dtype=None, reset_after_adc=True):
    """Final state of every voxel and offset row: [V, O, N] NumPy (working dtype)."""
    prm, m0, nc, V, code, npd = prepare(params, M_init, n_voxels, dtype)
    plan = plan_for(events, reset_after_adc)
    m, _ = _native.simulate(plan, prm, nc, code, m0, want_jac=False)
    return m.cpu().numpy().astype(npd, copy=False)


def simulate_trajectory(params: SimParams, events, M_init, n_voxels=None, dtype=None, reset_after_adc=True):
    """State after every event: [V, O, T, N]."""
    prm, m0, nc, V, code, npd = prepare(params, M_init, n_voxels, dtype)
    plan = plan_for(events, reset_after_adc, compress=False)
    out = _native.simulate_trajectory(plan, prm, nc, code, m0)
    return out.cpu().numpy().astype(npd, copy=False)


def simulate_with_jac(params: SimParams, events, M_init, slots, scales, jc, dM_init=None, n_voxels=None, dtype=None,
                      reset_after_adc=True, want_m=True, forward_jacobian=False):
    """Final states and d state[jc] / d x_p for packed slots (tangent scale applied).

    dM_init: [V, P, N] derivative of M_init per tangent, or None. forward_jacobian: forward-mode tangents
    even for single-event rows (the reverse kernel's result to rounding). Returns (m [V,O,N] or None,
    J [V,O,P]).
    """
    prm, m0, nc, V, code, npd = prepare(params, M_init, n_voxels, dtype)
    plan = plan_for(events, reset_after_adc)
    dm0 = None
    if dM_init is not None:
        N = 3 * (1 + nc) + 1
        dm0 = _native.device_array(np.broadcast_to(np.asarray(dM_init), (V, len(slots), N)).astype(npd),
                                   _torch_dtype(npd))
    m, J = _native.simulate(plan, prm, nc, code, m0, dm0=dm0, slots=slots, scales=scales, jc=jc, want_m=want_m,
                            forward_jacobian=forward_jacobian)
    m = None if m is None else m.cpu().numpy().astype(npd, copy=False)
    J = np.zeros((V, plan.n_offsets, 0), npd) if J is None else J.cpu().numpy().astype(npd, copy=False)
    return m, J
