"""Bounded Levenberg-Marquardt for ONE problem (src/bmc_jax/fitting/lm.py:20-72). Imports nothing of the package.

This is the generic form for an arbitrary Python residual callable, run on the host in float64 numpy
(fixed iterations, accept iff strictly better: lam x 0.5, else x 3; clip to the bounds; optional
`free` mask). The Jacobian is `resid_fn.jac(x)` when the callable provides one (the package's own
residuals do, computed by the GPU kernel's forward tangents); otherwise central differences.
Batched fits of the forward model do not come through here: `fit` / `fit_batch` run the whole loop
on the GPU (bmc_lm_fit).
"""
import numpy as np


def _fd_jacobian(f, x, r0):
    J = np.zeros((r0.size, x.size))
    for j in range(x.size):
        h = 1e-6 * max(abs(x[j]), 1.0)
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        J[:, j] = (np.asarray(f(xp), np.float64).reshape(-1) - np.asarray(f(xm), np.float64).reshape(-1)) / (2 * h)
    return J


def levenberg_marquardt(resid_fn, x0, lo, hi, n_iter=40, lam0=1e-2, use_fwd=False, return_trajectory=False,
                        free=None):
    """Damped Gauss-Newton with accept/reject (lm.py:43-61). use_fwd is accepted; the Jacobian is the same."""
    x = np.asarray(x0, np.float64).copy()
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    lam = float(lam0)
    P = x.size
    eye = np.eye(P)
    fr = None if free is None else np.asarray(free, np.float64)
    jac = getattr(resid_fn, "jac", None)
    X, C = [], []
    for _ in range(int(n_iter)):
        r = np.asarray(resid_fn(x), np.float64).reshape(-1)
        c = float(r @ r)
        J = np.asarray(jac(x), np.float64) if jac is not None else _fd_jacobian(resid_fn, x, r)
        if fr is not None:
            J = J * fr
        g = J.T @ r
        H = J.T @ J
        A = H + lam * (np.diag(np.diag(H)) + 1e-9 * eye)
        if fr is not None:
            A = A + np.diag(1.0 - fr)
        delta = np.linalg.solve(A, -g)
        xn = np.clip(x + delta, lo, hi)
        rn = np.asarray(resid_fn(xn), np.float64).reshape(-1)
        cn = float(rn @ rn)
        better = cn < c
        if better:
            x = xn
        lam = lam * 0.5 if better else lam * 3.0
        X.append(x.copy())
        C.append(c_out_select(better, cn, c))
    if return_trajectory:
        return x, np.array(X).reshape(-1, P), np.array(C)
    return x


def c_out_select(better, a, b):
    return a if better else b
