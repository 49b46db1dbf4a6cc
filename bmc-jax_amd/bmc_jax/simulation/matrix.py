"""The Bloch-McConnell system matrix of one event, as a host-side NumPy utility.

The kernels build this matrix themselves, entry by entry, inside the event scan (csrc/bmc_sim.hip);
this function exposes the same matrix to callers that want to inspect it, as the reference's
bmc_matrix does (src/bmc_jax/simulation/matrix.py:83-295): A (N x N) and C (N,) of dM/dt = A M + C,
with the Lorentzian (no 1/pi) and 101-point SuperLorentzian MT lineshapes of matrix.py:24-81.
"""
import numpy as np

from .params import SimParams, n_cest_of

_U = np.linspace(0.0, 1.0, 101)
_T = np.maximum(np.abs(3 * _U ** 2 - 1), 1e-6)


def _sl(delta, t2):
    a = delta * t2 / _T
    return float(np.sum(np.sqrt(2 / np.pi) * t2 / _T * np.exp(-2 * a ** 2)) * np.pi * 0.01)


def _chs(delta, t2, w0):
    px = [-300 - w0, -100 - w0, 100 + w0, 300 + w0]
    py = [_sl(x, t2) for x in px]
    d0, d1 = 30.0 * (py[1] - py[0]), 30.0 * (py[3] - py[2])
    c = abs((delta - px[1] + 1) / (px[2] - px[1] + 1))
    return ((2 * c ** 3 - 3 * c ** 2 + 1) * py[1] + (-2 * c ** 3 + 3 * c ** 2) * py[2] + (c ** 3 - 2 * c ** 2 + c) * d0
            + (c ** 3 - c ** 2) * d1)


def super_lorentzian(delta_omega: float, t2: float, w0: float) -> float:
    """SuperLorentzian lineshape with the cubic bridge inside |delta| < w0 (matrix.py:62-68)."""
    return _sl(delta_omega, t2) if abs(delta_omega) >= w0 else _chs(delta_omega, t2, w0)


def lorentzian(delta_omega: float, t2: float) -> float:
    """t2 / (1 + (delta t2)^2), in seconds (matrix.py:70-81)."""
    return t2 / (1 + (delta_omega * t2) ** 2)


def bmc_matrix(params: SimParams, rf_amp: float, rf_phase: float, rf_freq: float):
    """(A, C) of one event for one sample (matrix.py:83-295)."""
    g = lambda name: np.asarray(getattr(params, name), np.float64)
    k, f, r1c, r2c, dwc = g("cest_k"), g("cest_f"), g("cest_r1"), g("cest_r2"), g("cest_dw")
    r1w, r2w = float(g("water_r1")), float(g("water_r2"))
    kmt, fmt, r1mt, r2mt, dwmt = (float(g(n)) for n in ("mt_k", "mt_f", "mt_r1", "mt_r2", "mt_dw"))
    w1 = 2 * np.pi * rf_amp * float(g("rel_b1"))
    w1x, w1y = w1 * np.cos(rf_phase), w1 * np.sin(rf_phase)
    larmor = float(g("gamma")) * 1e6 * float(g("b0"))
    eff = 2 * np.pi * rf_freq - larmor * float(g("b0_inhom")) * 1e-6
    off_c = dwc * 1e-6 * larmor - eff
    off_mt = eff - dwmt * 1e-6 * larmor
    gmt = 0.0
    if fmt > 1e-9:
        ls = int(np.asarray(params.mt_lineshape))
        if ls == 1:
            gmt = lorentzian(off_mt, 1.0 / r2mt)
        elif ls == 2:
            gmt = super_lorentzian(off_mt, 1.0 / r2mt, larmor * 1e-6)
    w_mt = w1 ** 2 * gmt if fmt > 1e-9 else 0.0
    nc = n_cest_of(params)
    nt = 1 + nc
    N = 3 * nt + 1
    A = np.zeros((N, N))
    xw, yw, zw, zmt = 0, nt, 2 * nt, 3 * nt
    xc, yc, zc = np.arange(1, nt), np.arange(nt + 1, 2 * nt), np.arange(2 * nt + 1, 3 * nt)
    kwc = k * f
    ks = kwc.sum()
    A[xw, xw] = -r2w - ks; A[xw, yw] = eff; A[xw, zw] = -w1y
    A[yw, yw] = -r2w - ks; A[yw, xw] = -eff; A[yw, zw] = w1x
    A[zw, zw] = -r1w - ks - kmt * fmt; A[zw, yw] = -w1x; A[zw, xw] = w1y
    A[xc, xc] = -r2c - k; A[xc, yc] = -off_c; A[xc, zc] = -w1y
    A[yc, yc] = -r2c - k; A[yc, xc] = off_c; A[yc, zc] = w1x
    A[zc, zc] = -r1c - k; A[zc, xc] = w1y; A[zc, yc] = -w1x
    A[xw, xc] = k; A[yw, yc] = k; A[zw, zc] = k
    A[xc, xw] = kwc; A[yc, yw] = kwc; A[zc, zw] = kwc
    A[zmt, zmt] = -r1mt - kmt - w_mt; A[zmt, zw] = kmt * fmt; A[zw, zmt] = kmt
    C = np.zeros(N)
    C[zw] = r1w; C[zc] = r1c * f; C[zmt] = r1mt * fmt
    return A, C
