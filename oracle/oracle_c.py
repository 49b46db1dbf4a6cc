"""ctypes front end of the C oracle (oracle/bmc_oracle.c). TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker and by bench.py's cpu_baseline leg. Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbmc_oracle.so")
_lib = None

# packed record layout (include/bmc_hip.h)
SCALAR_SLOTS = {"water_r1": 0, "water_r2": 1, "water_f": 2, "mt_k": 3, "mt_f": 4, "mt_dw": 5, "mt_r1": 6,
                "mt_r2": 7, "mt_lineshape": 8, "b0": 9, "gamma": 10, "b0_inhom": 11, "rel_b1": 12}
POOL_FIELDS = ("cest_k", "cest_f", "cest_dw", "cest_r1", "cest_r2")


def slot_index(field, index, nc):
    if field in SCALAR_SLOTS:
        return SCALAR_SLOTS[field]
    return 13 + POOL_FIELDS.index(field) * nc + int(index)


def pack(p):
    """Oracle record (dict) -> packed float64 vector."""
    nc = len(p["cest_k"])
    out = np.zeros(13 + 5 * nc)
    for k, i in SCALAR_SLOTS.items():
        out[i] = float(p[k])
    for j, k in enumerate(POOL_FIELDS):
        out[13 + j * nc:13 + (j + 1) * nc] = np.asarray(p[k], float)
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"C oracle not built: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_simulate.argtypes = [ctypes.c_int] * 4 + [dp, ip, dp, dp, dp, dp, dp, dp, ctypes.c_int, ip,
                                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_simulate.restype = ctypes.c_int
        L.oracle_expm.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp, dp]
        L.oracle_expm.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def simulate(params_packed, nc, events, m0, slots=(), dm0=None, rel_h=1e-6, nan_rule=True, threads=None,
             want_m=True):
    """params_packed [V, 13+5nc]; events dict of (O, T); m0 [V, N] or [N]; slots: packed indices.

    Returns (m [V, O, N], dm [V, O, P, N] or None): final states and their derivatives w.r.t. the
    packed values (physical units).
    """
    L = lib()
    prm = np.ascontiguousarray(np.atleast_2d(params_packed), np.float64)
    V = prm.shape[0]
    N = 3 * (1 + nc) + 1
    typ = np.ascontiguousarray(events["type"], np.int32)
    O, T = typ.shape
    f = lambda k: np.ascontiguousarray(events[k], np.float64)
    dt, amp, ph, fr = f("dt"), f("rf_amp"), f("rf_phase"), f("rf_freq")
    m0 = np.ascontiguousarray(np.broadcast_to(np.asarray(m0, np.float64), (V, N)))
    P = len(slots)
    sl = np.ascontiguousarray(np.asarray(slots, np.int32).reshape(-1))
    d0 = None if dm0 is None else np.ascontiguousarray(np.broadcast_to(dm0, (V, P, N)), np.float64)
    m = np.zeros((V, O, N)) if want_m else None
    dm = np.zeros((V, O, P, N)) if P else None
    thr = threads if threads else (os.cpu_count() or 1)
    rc = L.oracle_simulate(V, nc, O, T, _dp(prm), _ip(typ), _dp(dt), _dp(amp), _dp(ph), _dp(fr), _dp(m0), _dp(d0),
                           P, _ip(sl) if P else None, rel_h, 1 if nan_rule else 0, thr, _dp(m), _dp(dm))
    if rc != 0:
        raise ValueError("oracle_simulate rejected its sizes")
    return m, dm


def expm(X, dX=None):
    """Batched float64 Pade-13 expm (Higham 2005) with optional tangents dX [B, P, n, n]."""
    L = lib()
    X = np.ascontiguousarray(np.asarray(X, np.float64))
    single = X.ndim == 2
    if single:
        X = X[None]
    B, n, _ = X.shape
    E = np.zeros_like(X)
    if dX is not None:
        dX = np.ascontiguousarray(np.asarray(dX, np.float64).reshape(B, -1, n, n))
        P = dX.shape[1]
        dE = np.zeros_like(dX)
    else:
        P, dE = 0, None
    rc = L.oracle_expm(B, n, _dp(X), P, _dp(dX), _dp(E), _dp(dE))
    if rc != 0:
        raise ValueError("oracle_expm rejected its sizes")
    if single:
        return (E[0], None if dE is None else dE[0])
    return E, dE
