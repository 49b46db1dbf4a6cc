"""ctypes binding of libbmc_hip.so (include/bmc_hip.h) and the device-memory plumbing around it.

PyTorch-ROCm is used here only for device memory and the current stream; every number is computed by
the library's HIP kernels. There is no CPU fallback: without the library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libbmc_hip.so")

BMC_F32, BMC_F64 = 0, 1
LM_EARLY_EXIT, LM_GRAPH, LM_START_CHECK, LM_FORWARD_JACOBIAN = 1, 2, 4, 8
SIM_FORWARD_JACOBIAN = 1
ABI_VERSION = 3
MAX_CEST = 4
MAX_TANGENTS = 16
EXPM_MAX_N = 24

# packed record layout (include/bmc_hip.h)
SCALAR_SLOTS = {"water_r1": 0, "water_r2": 1, "water_f": 2, "mt_k": 3, "mt_f": 4, "mt_dw": 5, "mt_r1": 6,
                "mt_r2": 7, "mt_lineshape": 8, "b0": 9, "gamma": 10, "b0_inhom": 11, "rel_b1": 12}
POOL_FIELDS = ("cest_k", "cest_f", "cest_dw", "cest_r1", "cest_r2")

_lib = None
_lock = threading.Lock()


class NativeError(RuntimeError):
    """A libbmc_hip call failed (message from bmc_last_error)."""


def slot_index(field: str, index, n_cest: int) -> int:
    """Packed-record position of a SimParams field (and pool index for cest_* fields)."""
    if field in SCALAR_SLOTS:
        return SCALAR_SLOTS[field]
    if field in POOL_FIELDS:
        if index is None or not 0 <= int(index) < n_cest:
            raise ValueError(f"{field}: pool index {index} outside 0..{n_cest - 1}")
        return 13 + POOL_FIELDS.index(field) * n_cest + int(index)
    raise ValueError(f"{field!r} is not a SimParams field")


def lib():
    """The loaded library. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"libbmc_hip.so is not built ({LIB_PATH}); run python bmc-jax_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip32, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.bmc_abi_version.restype = ctypes.c_int
        L.bmc_last_error.restype = ctypes.c_char_p
        L.bmc_device_count.restype = ctypes.c_int
        L.bmc_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, ip32, dp, dp, dp, dp, ctypes.c_int,
                                      ctypes.POINTER(vp)]
        L.bmc_plan_create_ex.argtypes = [ctypes.c_int, ctypes.c_int, ip32, dp, dp, dp, dp, ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(vp)]
        L.bmc_simulate_trajectory.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.bmc_plan_destroy.argtypes = [vp]
        L.bmc_plan_stats.argtypes = [vp, i64p, i64p, i64p]
        L.bmc_plan_program.argtypes = [vp, ctypes.c_int64, ip32, ip32]
        L.bmc_plan_compile.argtypes = [ctypes.c_int, ctypes.c_int, ip32, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int64,
                                       ip32, dp, ip32, i64p]
        L.bmc_simulate.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int, ip32, dp,
                                   ctypes.c_int, vp, vp, vp]
        L.bmc_simulate_ex.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int, ip32,
                                      dp, ctypes.c_int, vp, vp, ctypes.c_uint, vp]
        L.bmc_lm_fit.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, ip32, dp, dp, dp,
                                 ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_double, vp, ctypes.c_uint,
                                 ctypes.c_double, ctypes.c_double, vp, vp, vp, vp]
        L.bmc_lm_fit_noalloc.argtypes = L.bmc_lm_fit.argtypes
        L.bmc_tangents_per_pass.argtypes = [ctypes.c_int]
        L.bmc_tangents_per_pass.restype = ctypes.c_int
        L.bmc_adjoint_workspace_size.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p]
        L.bmc_simulate_adjoint.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int, ip32,
                                           dp, ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_int64, vp]
        L.bmc_expm.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, vp]
        for name in ("bmc_expm", "bmc_plan_create", "bmc_plan_create_ex", "bmc_plan_destroy", "bmc_plan_stats", "bmc_plan_program",
                     "bmc_plan_compile",
                     "bmc_simulate", "bmc_simulate_ex", "bmc_simulate_trajectory", "bmc_lm_fit", "bmc_lm_fit_noalloc", "bmc_adjoint_workspace_size",
                     "bmc_simulate_adjoint"):
            getattr(L, name).restype = ctypes.c_int
        if L.bmc_abi_version() != ABI_VERSION:
            raise NativeError(f"libbmc_hip.so has ABI version {L.bmc_abi_version()}, this binding needs {ABI_VERSION}: "
                              "rebuild with python bmc-jax_amd/build.py")
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().bmc_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({rc}): {msg}")


def torch_cuda():
    """torch with a visible GPU, or a loud error: the product path has no CPU fallback."""
    import torch
    if not torch.cuda.is_available():
        raise NativeError("no GPU visible: the BMC kernels run on HIP devices only")
    return torch


def stream_ptr():
    torch = torch_cuda()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _level(compress):
    """Plan compression level from a bool (True = 3) or an int 0..3."""
    if compress is True:
        return 3
    if compress is False or compress is None:
        return 0
    lv = int(compress)
    if lv not in (0, 1, 2, 3):
        raise ValueError(f"compression level must be 0, 1, 2 or 3, got {compress}")
    return lv


def compile_program(type_, dt, rf_amp, rf_phase, rf_freq, compress=3):
    """The row programs of an event table, compiled on the host only (bmc_plan_compile; needs no GPU):
    (ops [n_ops, 5] int32 = (kind, first event, length, count, half), rot [n_ops, 2] = (cos d, sin d),
    row_ptr [O + 1] int32). Kinds: 0 SEQ, 1 BLOCK, 2 KEEP, 3 REUSE. half h > 0: the op's first 2h events
    form a palindrome at one RF phase (its cos, sin in rot), chained from their first half."""
    L = lib()
    t = np.ascontiguousarray(np.asarray(type_), dtype=np.int32)
    if t.ndim != 2:
        raise ValueError(f"events must be (n_offsets, n_events), got shape {t.shape}")
    f = [np.ascontiguousarray(np.asarray(a, np.float64)) for a in (dt, rf_amp, rf_phase, rf_freq)]
    O, T = t.shape
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    n = ctypes.c_int64()
    args = [O, T, ip(t)] + [_dp(a) for a in f] + [_level(compress)]
    check(L.bmc_plan_compile(*args, 0, None, None, None, ctypes.byref(n)), "bmc_plan_compile")
    ops = np.zeros((max(n.value, 1), 5), np.int32)
    rot = np.zeros((max(n.value, 1), 2), np.float64)
    rp = np.zeros(O + 1, np.int32)
    check(L.bmc_plan_compile(*args, max(n.value, 1), ip(ops), _dp(rot), ip(rp), ctypes.byref(n)), "bmc_plan_compile")
    return ops[:n.value], rot[:n.value], rp


class Plan:
    """A sequence compiled for the device (bmc_plan_create): events [O, T] and the row programs."""

    def __init__(self, type_, dt, rf_amp, rf_phase, rf_freq, reset_after_adc=True, compress=3):
        """compress: 0 every event its own step; 1 runs of identical events chained; 2 also pulses that
        recur with a shifted RF phase kept and reused; 3 (True) also palindromic runs built from their
        first half (bmc_plan_create_ex)."""
        L = lib()
        torch_cuda()
        t = np.ascontiguousarray(np.asarray(type_), dtype=np.int32)
        if t.ndim != 2:
            raise ValueError(f"events must be (n_offsets, n_events), got shape {t.shape}")
        f = lambda a: np.ascontiguousarray(np.asarray(a, np.float64))
        self._arrays = (t, f(dt), f(rf_amp), f(rf_phase), f(rf_freq))
        for a in self._arrays[1:]:
            if a.shape != t.shape:
                raise ValueError("event fields disagree in shape")
        self.n_offsets, self.n_events = t.shape
        h = ctypes.c_void_p()
        check(L.bmc_plan_create_ex(self.n_offsets, self.n_events, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   *[_dp(a) for a in self._arrays[1:]], 1 if reset_after_adc else 0,
                                   _level(compress), ctypes.byref(h)), "bmc_plan_create_ex")
        self.compressed = bool(compress)
        self.level = _level(compress)
        self._h = h
        self._arrays = None

    @property
    def handle(self):
        return self._h

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().bmc_plan_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "bmc_plan_stats")
        return dict(n_ops=a.value, n_built=b.value, n_applied=c.value)

    def program(self):
        """(ops [n_ops, 5] int32, row_ptr [O + 1] int32): the compiled row programs."""
        n = self.stats()["n_ops"]
        ops = np.zeros((max(n, 1), 5), np.int32)
        rp = np.zeros(self.n_offsets + 1, np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        check(lib().bmc_plan_program(self._h, max(n, 1), ip(ops), ip(rp)), "bmc_plan_program")
        return ops[:n], rp

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            try:
                _lib.bmc_plan_destroy(h)
            except Exception:
                pass
            self._h = None


def device_array(x, dtype):
    """numpy -> contiguous device tensor of the given torch dtype."""
    torch = torch_cuda()
    t = torch.as_tensor(np.ascontiguousarray(x))
    return t.to(device="cuda", dtype=dtype).contiguous()


def expm(a):
    """bmc_expm: exp of every trailing n x n matrix of `a` (numpy, float32 or float64, any leading axes),
    computed on the GPU in a's precision; returned as numpy of a's dtype and shape."""
    torch = torch_cuda()
    a = np.asarray(a)
    if a.dtype not in (np.float32, np.float64):
        raise ValueError(f"expm takes float32 or float64 matrices, not {a.dtype}")
    if a.ndim < 2 or a.shape[-1] != a.shape[-2]:
        raise ValueError(f"expm takes square matrices in the last two axes, not shape {a.shape}")
    n = a.shape[-1]
    batch = int(np.prod(a.shape[:-2], dtype=np.int64))
    f64 = a.dtype == np.float64
    t = device_array(a.reshape(batch, n, n), torch.float64 if f64 else torch.float32)
    out = torch.empty_like(t)
    check(lib().bmc_expm(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(out.data_ptr()), n, batch,
                         BMC_F64 if f64 else BMC_F32, stream_ptr()), "bmc_expm")
    return out.cpu().numpy().reshape(a.shape)


def simulate(plan: Plan, params, n_cest: int, dtype_code: int, m0, dm0=None, slots=(), scales=(), jc=0,
             want_m=True, want_jac=True, forward_jacobian=False):
    """bmc_simulate_ex on device tensors. params [V, S], m0 [V, N] torch tensors of the working dtype.
    forward_jacobian: forward-mode tangents even where the reverse-accumulation kernel applies.

    Returns (m [V, O, N] or None, jac [V, O, P] or None) as device tensors.
    """
    torch = torch_cuda()
    L = lib()
    V = int(params.shape[0])
    N = 3 * (1 + n_cest) + 1
    P = len(slots)
    tdt = torch.float64 if dtype_code == BMC_F64 else torch.float32
    m_out = torch.empty((V, plan.n_offsets, N), dtype=tdt, device="cuda") if want_m else None
    jac = torch.empty((V, plan.n_offsets, P), dtype=tdt, device="cuda") if (want_jac and P) else None
    sl = (ctypes.c_int32 * max(P, 1))(*[int(s) for s in slots])
    sc = (ctypes.c_double * max(P, 1))(*[float(s) for s in scales])
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    flags = SIM_FORWARD_JACOBIAN if forward_jacobian else 0
    check(L.bmc_simulate_ex(plan.handle, V, n_cest, dtype_code, ptr(params), ptr(m0), ptr(dm0), P, sl, sc, jc,
                            ptr(m_out), ptr(jac), flags, stream_ptr()), "bmc_simulate_ex")
    return m_out, jac


def simulate_trajectory(plan: Plan, params, n_cest: int, dtype_code: int, m0):
    """bmc_simulate_trajectory: the state after every event [V, O, T, N] (device tensor)."""
    torch = torch_cuda()
    V = int(params.shape[0])
    N = 3 * (1 + n_cest) + 1
    tdt = torch.float64 if dtype_code == BMC_F64 else torch.float32
    out = torch.empty((V, plan.n_offsets, plan.n_events, N), dtype=tdt, device="cuda")
    check(lib().bmc_simulate_trajectory(plan.handle, V, n_cest, dtype_code, ctypes.c_void_p(params.data_ptr()),
                                        ctypes.c_void_p(m0.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                        stream_ptr()), "bmc_simulate_trajectory")
    return out


def lm_fit(plan: Plan, params, n_cest: int, dtype_code: int, m0, slots, scales, lo, hi, jc, z_meas, x0, n_iter,
           lam0=1e-2, free=None, flags=0, early_ftol=0.0, early_rms=0.0, want_traj=False, want_cost=False,
           noalloc=False):
    """bmc_lm_fit on device tensors. x0 [V, P] float64 (scaled). Returns (x, cost_traj, x_traj, n_done);
    cost_traj [V, n_iter] comes with want_traj or want_cost, x_traj only with want_traj. noalloc: the
    allocation-free entry point bmc_lm_fit_noalloc (its workspace is the library's fixed device array;
    it does not synchronise the stream)."""
    torch = torch_cuda()
    V, P = int(x0.shape[0]), len(slots)
    x = x0.to(device="cuda", dtype=torch.float64).contiguous().clone()
    ctraj = torch.empty((V, n_iter), dtype=torch.float64, device="cuda") if (want_traj or want_cost) else None
    xtraj = torch.empty((V, n_iter, P), dtype=torch.float64, device="cuda") if want_traj else None
    ndone = torch.empty((V,), dtype=torch.int32, device="cuda")
    arr = lambda ct, vals: (ct * max(len(vals), 1))(*vals)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    fr = None if free is None else free.to(device="cuda", dtype=torch.float64).contiguous()
    entry = lib().bmc_lm_fit_noalloc if noalloc else lib().bmc_lm_fit
    check(entry(plan.handle, V, n_cest, dtype_code, ptr(params), ptr(m0), P,
                           arr(ctypes.c_int32, [int(s) for s in slots]), arr(ctypes.c_double, [float(s) for s in scales]),
                           arr(ctypes.c_double, [float(v) for v in lo]), arr(ctypes.c_double, [float(v) for v in hi]),
                           jc, ptr(z_meas), ptr(x), int(n_iter), float(lam0), ptr(fr), int(flags), float(early_ftol),
                           float(early_rms), ptr(ctraj), ptr(xtraj), ptr(ndone), stream_ptr()),
          "bmc_lm_fit_noalloc" if noalloc else "bmc_lm_fit")
    return x, ctraj, xtraj, ndone


def simulate_adjoint(plan: Plan, params, n_cest: int, dtype_code: int, m0, slots=(), scales=(), jc=0, z_cot=None,
                     jac_cot=None, want_amp=True, want_freq=True, dm0=None):
    """bmc_simulate_adjoint: (dL/d rf_amp [O, T], dL/d rf_freq [O, T]) float64 device tensors (None where not
    wanted) for cotangents z_cot [V, O] on Z and jac_cot [V, O, P] on the Jacobian of simulate()."""
    torch = torch_cuda()
    L = lib()
    V = int(params.shape[0])
    P = len(slots)
    nbytes = ctypes.c_int64()
    check(L.bmc_adjoint_workspace_size(plan.handle, V, n_cest, P, ctypes.byref(nbytes)), "bmc_adjoint_workspace_size")
    ws = torch.empty(max(int(nbytes.value), 8), dtype=torch.uint8, device="cuda")
    shape = (plan.n_offsets, plan.n_events)
    ga = torch.zeros(shape, dtype=torch.float64, device="cuda") if want_amp else None
    gf = torch.zeros(shape, dtype=torch.float64, device="cuda") if want_freq else None
    f64 = lambda t: None if t is None else t.to(device="cuda", dtype=torch.float64).contiguous()
    zc, jcot = f64(z_cot), f64(jac_cot)
    sl = (ctypes.c_int32 * max(P, 1))(*[int(s) for s in slots])
    sc = (ctypes.c_double * max(P, 1))(*[float(s) for s in scales])
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    check(L.bmc_simulate_adjoint(plan.handle, V, n_cest, dtype_code, ptr(params), ptr(m0), ptr(dm0), P, sl, sc, jc,
                                 ptr(zc), ptr(jcot), ptr(ga), ptr(gf), ptr(ws), int(nbytes.value), stream_ptr()),
          "bmc_simulate_adjoint")
    return ga, gf
