"""Differentiating the Cramer-Rao design cost through the sequence with the adjoint kernel.

The reference differentiates `cost(d)` with jax.value_and_grad, reverse mode over the forward-mode
Jacobian of the scan (src/bmc_jax/design/optimize.py:80-93). Here the same chain is a torch graph:
`build(d)` (caller code, torch operations) -> events whose rf_amp / rf_freq are tensors ->
ScanJacobian (bmc_simulate forward, bmc_simulate_adjoint backward) -> Fisher matrix, inverse and
cost in float64 torch.

ScanJacobian maps (rf_amp [O, T], rf_freq [O, T]) to (Z [V, O], J [V, O, P]) for V parameter sets
sharing the sequence, J = dZ/dx_p * scale_p. Its backward returns dL/d rf_amp and dL/d rf_freq.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import _engine, _native
from ..fitting.params import write_params
from ..fitting.residual import spec_slots
from ..simulation.params import n_cest_of
from ..simulation.voxels import tile_params


@dataclass
class ScanProblem:
    """Everything about a scan except the RF amplitudes and frequencies."""
    type: np.ndarray       # [O, T] int32
    dt: np.ndarray         # [O, T] float64
    rf_phase: np.ndarray   # [O, T] float64
    params: object         # device tensor [V, 13 + 5 nc], working dtype
    m0: object             # device tensor [V, N]
    n_cest: int
    dtype_code: int
    slots: list
    scales: list
    jc: int
    reset_after_adc: bool = True


def _host(x):
    if isinstance(x, torch.Tensor):
        return x.detach().to("cpu", torch.float64).numpy()
    return np.asarray(x, np.float64)


class ScanJacobian(torch.autograd.Function):
    """(rf_amp, rf_freq) -> (Z, J) on the GPU; the backward is the adjoint kernel."""

    @staticmethod
    def forward(ctx, rf_amp, rf_freq, prob: ScanProblem):
        plan = _native.Plan(prob.type, prob.dt, _host(rf_amp), prob.rf_phase, _host(rf_freq), prob.reset_after_adc)
        m, J = _native.simulate(plan, prob.params, prob.n_cest, prob.dtype_code, prob.m0, slots=prob.slots,
                                scales=prob.scales, jc=prob.jc, want_m=True, want_jac=bool(prob.slots))
        ctx.plan, ctx.prob = plan, prob
        ctx.amp_dtype, ctx.freq_dtype = rf_amp.dtype, rf_freq.dtype
        ctx.amp_device, ctx.freq_device = rf_amp.device, rf_freq.device
        Z = m[..., prob.jc].contiguous()
        if J is None:
            J = torch.zeros(Z.shape + (0,), dtype=Z.dtype, device=Z.device)
        return Z, J

    @staticmethod
    def backward(ctx, gZ, gJ):
        prob = ctx.prob
        want_a, want_f = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (want_a or want_f):
            return None, None, None
        ga, gf = _native.simulate_adjoint(ctx.plan, prob.params, prob.n_cest, prob.dtype_code, prob.m0,
                                          slots=prob.slots, scales=prob.scales, jc=prob.jc, z_cot=gZ,
                                          jac_cot=gJ if prob.slots else None, want_amp=want_a, want_freq=want_f)
        to = lambda g, dt, dev: None if g is None else g.to(device=dev, dtype=dt)
        return to(ga, ctx.amp_dtype, ctx.amp_device), to(gf, ctx.freq_dtype, ctx.freq_device), None


def _events_tensors(events):
    """(rf_amp, rf_freq) as float64 tensors (kept as given when they already are tensors)."""
    f = lambda x: x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, np.float64))
    return f(events.rf_amp), f(events.rf_freq)


def is_differentiable(events) -> bool:
    return any(isinstance(x, torch.Tensor) and x.requires_grad for x in (events.rf_amp, events.rf_freq))


def scan_jacobian(experiment, spec, truths, dtype=None):
    """(Z [V, O], J [V, O, P]) in physical units for every row of truths [V, P] (physical), with the
    experiment's sequence; differentiable w.r.t. its rf_amp / rf_freq tensors."""
    truths = np.atleast_2d(np.asarray(truths, np.float64))
    V = truths.shape[0]
    base = experiment.params
    nc = n_cest_of(base)
    _engine.check_pools(nc)
    code, npd = _engine.dtype_code(dtype)
    tdt = torch.float64 if npd == np.float64 else torch.float32
    pb = write_params(tile_params(base, V), truths / spec.scales, spec)
    prm, m0, _, _, _, _ = _engine.prepare(pb, np.broadcast_to(np.asarray(experiment.m0), (V, np.shape(experiment.m0)[-1])),
                                          n_voxels=V, dtype=npd)
    ev = experiment.events
    prob = ScanProblem(type=np.asarray(ev.type, np.int32), dt=_host(ev.dt), rf_phase=_host(ev.rf_phase),
                       params=prm.to(tdt), m0=m0.to(tdt), n_cest=nc, dtype_code=code,
                       slots=spec_slots(spec, nc), scales=[1.0] * spec.n, jc=int(experiment.mz_loc))
    amp, freq = _events_tensors(ev)
    return ScanJacobian.apply(amp, freq, prob)


def design_cost(experiment, spec, sigma, truths, target_idx, dtype=None):
    """mean over truths of sum_targets diag(F^-1)_j / x_j^2, F = J^T J / sigma^2 (optimize.py:80-87), as a
    float64 torch scalar that carries the gradient w.r.t. the experiment's RF tensors."""
    truths = np.atleast_2d(np.asarray(truths, np.float64))
    _, J = scan_jacobian(experiment, spec, truths, dtype)
    J = J.to(torch.float64)
    F = torch.matmul(J.transpose(1, 2).contiguous(), J) / float(sigma) ** 2
    var = torch.diagonal(torch.linalg.inv(F), dim1=1, dim2=2)[:, target_idx]
    x = torch.as_tensor(truths[:, target_idx], dtype=torch.float64, device=var.device)
    return torch.mean(torch.sum(var / x ** 2, dim=1))
