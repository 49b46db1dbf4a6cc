"""A protocol with a lower Cramer-Rao bound, by gradient descent (src/bmc_jax/design/optimize.py:31-113).

cost(d) = mean over truths of sum over targets of var_j / x_j^2, with d = lo + (hi - lo) sigmoid(u) and
Adam on u. How the gradient of the cost with respect to d is taken depends on `build(d)`:
  * `build` written with torch operations, returning an Experiment whose events' rf_amp / rf_freq are
    tensors that depend on d: reverse mode through torch and the adjoint kernel (design/adjoint.py),
    one forward and one adjoint launch per step whatever the length of d -- the reference's
    value_and_grad (optimize.py:90-93);
  * any other `build` (NumPy): central differences of the GPU-evaluated cost, two cost evaluations
    per design entry.
"""
from dataclasses import dataclass, replace
from typing import Callable, Optional, Sequence

import numpy as np

from .crlb import crlb, fisher_information


@dataclass(frozen=True)
class DesignResult:
    design: np.ndarray
    start: np.ndarray
    cost: np.ndarray
    targets: tuple
    truths: np.ndarray
    rel_sd_start: np.ndarray
    rel_sd_end: np.ndarray


def _numpy_experiment(exp):
    """The experiment with its event tensors (if any) as NumPy arrays."""
    ev = exp.events
    conv = lambda x: x.detach().cpu().numpy() if hasattr(x, "detach") else x
    return replace(exp, events=type(ev)(*[conv(f) for f in ev]))


def _build_numpy(build, d):
    """build(d) for the bound reports: tried with a NumPy d first, then with a float64 tensor."""
    try:
        return _numpy_experiment(build(np.asarray(d, float)))
    except Exception:
        import torch
        return _numpy_experiment(build(torch.as_tensor(np.asarray(d, float))))


def optimize_design(build: Callable, spec, d0, lo, hi, sigma: float, truths, targets: Optional[Sequence[str]] = None,
                    n_steps: int = 200, lr: float = 0.05, callback: Optional[Callable] = None,
                    fd_step: float = 1e-4, gradient: str = "auto") -> DesignResult:
    """Lower the bound of the target parameters by changing the design (optimize.py:43-113).

    gradient: "adjoint" (build must use torch operations), "fd" (central differences), or "auto"
    (adjoint when build(d) for a tensor d returns differentiable event tensors, else differences).
    """
    names = [s.name for s in spec.slots]
    targets = tuple(names if targets is None else targets)
    idx = np.asarray([names.index(t) for t in targets])
    truths = np.atleast_2d(np.asarray(truths, float))
    d0, lo, hi = (np.asarray(a, float) for a in (d0, lo, hi))
    if not (d0.shape == lo.shape == hi.shape) or np.any(d0 <= lo) or np.any(d0 >= hi):
        raise ValueError(f"the start must lie strictly inside its box, lo < d0 < hi elementwise, with d0, lo and "
                         f"hi of one shape (got {d0.shape}, {lo.shape}, {hi.shape})")

    def cost_of(d):
        exp = _numpy_experiment(build(d))
        total = 0.0
        for x in truths:
            F, _ = fisher_information(exp, spec, x, sigma)
            var = np.diag(np.linalg.inv(F))[idx]
            total += float(np.sum(var / np.asarray(x)[idx] ** 2))
        return total / len(truths)

    sig = lambda u: 1.0 / (1.0 + np.exp(-u))
    to_d = lambda u: lo + (hi - lo) * sig(u)
    p0 = (d0 - lo) / (hi - lo)
    u = np.log(p0 / (1.0 - p0))

    def value_and_grad_fd(u):
        c = cost_of(to_d(u))
        g = np.zeros_like(u)
        for j in range(u.size):
            e = np.zeros_like(u)
            e[j] = fd_step
            g[j] = (cost_of(to_d(u + e)) - cost_of(to_d(u - e))) / (2 * fd_step)
        return c, g

    def value_and_grad_adjoint(u):
        import torch
        from .adjoint import design_cost
        ut = torch.tensor(u, dtype=torch.float64, requires_grad=True)
        lo_t, hi_t = torch.as_tensor(lo), torch.as_tensor(hi)
        exp = build(lo_t + (hi_t - lo_t) * torch.sigmoid(ut))
        c = design_cost(exp, spec, sigma, truths, idx)
        c.backward()
        return float(c.detach()), ut.grad.numpy().copy()

    if gradient not in ("auto", "adjoint", "fd"):
        raise ValueError(f"gradient must be 'auto', 'adjoint' or 'fd', got {gradient!r}")
    use_adjoint = gradient == "adjoint"
    if gradient == "auto":
        try:
            import torch
            from .adjoint import is_differentiable
            probe = build(torch.tensor(d0, dtype=torch.float64, requires_grad=True))
            use_adjoint = is_differentiable(probe.events)
        except Exception:
            use_adjoint = False
    value_and_grad = value_and_grad_adjoint if use_adjoint else value_and_grad_fd

    m, v, b1, b2, eps = np.zeros_like(u), np.zeros_like(u), 0.9, 0.999, 1e-8
    history = []
    for step in range(n_steps):
        c, g = value_and_grad(u)
        if not (np.isfinite(c) and np.all(np.isfinite(g))):
            raise FloatingPointError(f"step {step}: the design cost or its gradient is NaN or infinite (a smaller "
                                     "lr or float64 may help)")
        history.append(float(c))
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g ** 2
        u = u - lr * (m / (1 - b1 ** (step + 1))) / (np.sqrt(v / (1 - b2 ** (step + 1))) + eps)
        if callback is not None:
            callback(step, float(c), np.asarray(to_d(u)))
    design = np.asarray(to_d(u), float)
    if use_adjoint:
        import torch
        from .adjoint import design_cost
        with torch.no_grad():
            history.append(float(design_cost(build(torch.as_tensor(design)), spec, sigma, truths, idx)))
    else:
        history.append(float(cost_of(design)))

    def rel(d):
        exp = _build_numpy(build, d)
        return np.array([[crlb(exp, spec, sigma, x).rel_sd[names.index(t)] for t in targets] for x in truths])
    return DesignResult(design, d0, np.asarray(history), targets, truths, rel(d0), rel(design))
