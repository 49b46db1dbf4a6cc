"""Uneven batched fitting: many independent problems of different shapes in one GPU LM.

The reference fits problems of different shapes (pool count, rungs x offsets, free values) one `fit`
call at a time (src/bmc_jax/fitting/independent.py:226-298), because vmap needs equal shapes. Here
`fit_uneven` fits them together: every problem's rows go into one plan (front-padded with zero-length
delays, an exact no-op: simulation/events.py:34-91), and each LM sweep over all problems is one
persistent launch (libbmc_hip's bmc_ubatch_*). The launch mapping -- tangents per pass, persistent grid
size, item order -- is chosen by `autotune`, which times the candidates on the batch itself; the
results do not depend on the choice.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .. import _engine, _native
from ..simulation.events import SimEvent, concat_events, pad_events
from ..simulation.params import SimParams, n_cest_of, pack
from .params import FitSpec
from .residual import spec_slots


@dataclass
class UnevenProblem:
    """One fit problem: a sample, its sequence (O rows), its start state, what to fit, its data."""
    params: SimParams
    events: SimEvent
    m0: np.ndarray
    spec: FitSpec
    z_meas: np.ndarray
    x0_phys: Optional[np.ndarray] = None
    mz_loc: Optional[int] = None


_BOUND = False


def _bind(L):
    global _BOUND
    if _BOUND:
        return
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
    L.bmc_ubatch_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip, ip, ip, ip, ip, dp, dp, dp,
                                    ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.bmc_ubatch_configure.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.bmc_ubatch_sizes.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int64)]
    L.bmc_ubatch_simulate.argtypes = [vp, vp, vp, vp, vp]
    L.bmc_ubatch_simulate_ex.argtypes = [vp, vp, vp, vp, ctypes.c_uint, vp]
    L.bmc_ubatch_lm_fit.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_uint, ctypes.c_double,
                                    ctypes.c_double, vp, vp, vp]
    L.bmc_ubatch_destroy.argtypes = [vp]
    for n in ("bmc_ubatch_create", "bmc_ubatch_configure", "bmc_ubatch_sizes", "bmc_ubatch_simulate",
              "bmc_ubatch_simulate_ex",
              "bmc_ubatch_lm_fit", "bmc_ubatch_destroy"):
        getattr(L, n).restype = ctypes.c_int
    _BOUND = True


class UnevenBatch:
    """Device form of a list of UnevenProblem (bmc_ubatch_create)."""

    def __init__(self, problems: Sequence[UnevenProblem], dtype=None):
        torch = _native.torch_cuda()
        L = _native.lib()
        _bind(L)
        problems = list(problems)
        if not problems:
            raise ValueError("no problems")
        self.problems = problems
        self.code, self.npd = _engine.dtype_code(dtype)
        tdt = torch.float64 if self.npd == np.float64 else torch.float32
        n = len(problems)
        width = max(int(np.shape(p.events.dt)[1]) for p in problems)
        rows, row0, n_rows, ncs, nps = [], [], [], [], []
        slots, scales, lo, hi = [], [], [], []
        by_cls = {c: ([], []) for c in range(4)}
        r = 0
        for i, p in enumerate(problems):
            nc = n_cest_of(p.params)
            if not 1 <= nc <= 4:
                raise ValueError(f"problem {i}: {nc} CEST pools; the uneven batch takes 1..4")
            O = int(np.shape(p.events.dt)[0])
            z = np.asarray(p.z_meas)
            if z.shape != (O,):
                raise ValueError(f"problem {i}: z_meas has shape {z.shape}, its sequence has {O} rows")
            p.spec.require_independent("fit_uneven")
            mz = 2 * (1 + nc) if p.mz_loc is None else int(p.mz_loc)
            if mz != 2 * (1 + nc):
                raise ValueError(f"problem {i}: mz_loc {mz} is not water Mz ({2 * (1 + nc)})")
            rows.append(pad_events(p.events, width))
            row0.append(r)
            n_rows.append(O)
            r += O
            ncs.append(nc)
            nps.append(p.spec.n)
            slots += spec_slots(p.spec, nc)
            scales += list(p.spec.scales)
            lo += list(p.spec.lo_phys / p.spec.scales)
            hi += list(p.spec.hi_phys / p.spec.scales)
            by_cls[nc - 1][0].append(pack(p.params, 1, dtype=self.npd)[0])
            by_cls[nc - 1][1].append(np.asarray(p.m0, self.npd))
        self.R, self.n = r, n
        self.row0, self.n_rows, self.n_params = np.array(row0), np.array(n_rows), np.array(nps)
        self.x_off = np.concatenate([[0], np.cumsum(nps)])
        self.scales = np.array(scales)
        self.plan_events = concat_events(rows)
        # level 1: the persistent sweep holds no LDS slot for pulses kept across a row
        self.plan = _engine.plan_for(self.plan_events, compress=1)
        self._keep = []
        prm_ptrs = (ctypes.c_void_p * 4)()
        m0_ptrs = (ctypes.c_void_p * 4)()
        for c in range(4):
            if by_cls[c][0]:
                tp = _native.device_array(np.stack(by_cls[c][0]), tdt)
                tm = _native.device_array(np.stack(by_cls[c][1]), tdt)
                self._keep += [tp, tm]
                prm_ptrs[c] = tp.data_ptr()
                m0_ptrs[c] = tm.data_ptr()
        ia = lambda v: (ctypes.c_int32 * len(v))(*[int(x) for x in v])
        da = lambda v: (ctypes.c_double * len(v))(*[float(x) for x in v])
        h = ctypes.c_void_p()
        _native.check(L.bmc_ubatch_create(self.plan.handle, n, self.code, ia(ncs), ia(nps), ia(row0), ia(n_rows),
                                          ia(slots), da(scales), da(lo), da(hi), prm_ptrs, m0_ptrs, ctypes.byref(h)),
                      "bmc_ubatch_create")
        self._h = h
        nr, nv, nj = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _native.check(L.bmc_ubatch_sizes(h, ctypes.byref(nr), ctypes.byref(nv), ctypes.byref(nj)), "bmc_ubatch_sizes")
        self.n_values, self.n_jacobian = nv.value, nj.value
        self.config = (1, 1024, 1)
        self.z_meas = _native.device_array(np.concatenate([np.asarray(p.z_meas) for p in problems]).astype(self.npd),
                                           tdt)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _native._lib is not None:
            _native._lib.bmc_ubatch_destroy(h)
            self._h = None

    def configure(self, tangents_per_pass=1, grid=1024, order=1):
        _native.check(_native.lib().bmc_ubatch_configure(self._h, int(tangents_per_pass), int(grid), int(order)),
                      "bmc_ubatch_configure")
        self.config = (int(tangents_per_pass), int(grid), int(order))

    def x0_scaled(self):
        return np.concatenate([p.spec.to_scaled(p.spec.x0_phys if p.x0_phys is None else p.x0_phys)
                               for p in self.problems])

    def simulate(self, x_scaled, jacobian=True, forward_jacobian=False):
        """(z [R], J ragged [sum n_rows P] or None) as float64 device tensors at x (scaled, concatenated).
        forward_jacobian: the forward-mode sweep even for single-event rows."""
        torch = _native.torch_cuda()
        x = torch.as_tensor(np.asarray(x_scaled, np.float64), device="cuda")
        z = torch.empty(self.R, dtype=torch.float64, device="cuda")
        J = torch.empty(max(self.n_jacobian, 1), dtype=torch.float64, device="cuda") if jacobian else None
        flags = _native.SIM_FORWARD_JACOBIAN if forward_jacobian else 0
        _native.check(_native.lib().bmc_ubatch_simulate_ex(self._h, ctypes.c_void_p(x.data_ptr()),
                                                            ctypes.c_void_p(z.data_ptr()),
                                                            ctypes.c_void_p(J.data_ptr()) if J is not None else None,
                                                            flags, _native.stream_ptr()), "bmc_ubatch_simulate_ex")
        return z, J

    def jacobian_of(self, J, i):
        """Problem i's block [n_rows, P] of the ragged Jacobian (NumPy)."""
        Jh = J.cpu().numpy() if hasattr(J, "cpu") else np.asarray(J)
        off = int(np.sum(self.n_rows[:i] * self.n_params[:i]))
        return Jh[off:off + self.n_rows[i] * self.n_params[i]].reshape(self.n_rows[i], self.n_params[i])

    def lm_fit(self, x0_scaled=None, n_iter=30, lam0=1e-2, early_exit=False, early_ftol=1e-10, early_rms=0.0,
               graph=True, forward_jacobian=False):
        """bmc_ubatch_lm_fit. Returns (x scaled [sum P], final cost [n], iterations run [n]) as NumPy."""
        torch = _native.torch_cuda()
        x = torch.as_tensor(np.asarray(self.x0_scaled() if x0_scaled is None else x0_scaled, np.float64),
                            device="cuda").clone()
        cost = torch.empty(self.n, dtype=torch.float64, device="cuda")
        nd = torch.empty(self.n, dtype=torch.int32, device="cuda")
        flags = ((_native.LM_EARLY_EXIT if early_exit else 0) | (_native.LM_GRAPH if graph else 0) |
                 (_native.LM_FORWARD_JACOBIAN if forward_jacobian else 0))
        _native.check(_native.lib().bmc_ubatch_lm_fit(self._h, ctypes.c_void_p(self.z_meas.data_ptr()),
                                                       ctypes.c_void_p(x.data_ptr()), int(n_iter), float(lam0),
                                                       flags, float(early_ftol), float(early_rms),
                                                       ctypes.c_void_p(cost.data_ptr()), ctypes.c_void_p(nd.data_ptr()),
                                                       _native.stream_ptr()), "bmc_ubatch_lm_fit")
        torch.cuda.synchronize()
        return x.cpu().numpy(), cost.cpu().numpy(), nd.cpu().numpy()

    def split(self, x_scaled):
        """Per-problem physical values from the concatenated scaled vector."""
        return [np.asarray(x_scaled[self.x_off[i]:self.x_off[i + 1]]) * self.problems[i].spec.scales
                for i in range(self.n)]

    def signature(self):
        """What the best launch mapping depends on: problem-class histogram, rows and values."""
        cls = np.bincount(np.array([n_cest_of(p.params) for p in self.problems]), minlength=5)
        return (tuple(int(c) for c in cls), int(self.R), int(self.n_values), int(self.n_jacobian))


_TUNED = {}

CANDIDATES = [(pt, grid, order) for pt in (1, 2) for grid in (256, 512, 1024, 2048) for order in (0, 1)]


def autotune(batch: UnevenBatch, candidates=CANDIDATES, repeats=2, log=None):
    """Time the tangent sweep (the costliest launch of an iteration) for every candidate mapping on this
    batch, keep the fastest, and remember it for batches of the same signature. Returns
    (best config, {config: ms})."""
    torch = _native.torch_cuda()
    key = batch.signature()
    if key in _TUNED:
        batch.configure(*_TUNED[key][0])
        return _TUNED[key]
    x = batch.x0_scaled()
    times = {}
    for cfg in candidates:
        batch.configure(*cfg)
        batch.simulate(x)                      # warm
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            batch.simulate(x)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        times[cfg] = best
        if log:
            log(f"autotune {cfg}: {best:.2f} ms")
    cfg = min(times, key=times.get)
    batch.configure(*cfg)
    _TUNED[key] = (cfg, times)
    return _TUNED[key]


def fit_uneven(problems: Sequence[UnevenProblem], n_iter: int = 30, dtype=None, tune: bool = True,
               early_exit: bool = False, graph: bool = True, lam0: float = 1e-2) -> List[np.ndarray]:
    """Fit every problem with the bounded LM of `fit` (lm.py:20-67), all problems in one GPU batch.
    Returns the physical values of each problem."""
    batch = UnevenBatch(problems, dtype)
    if tune:
        autotune(batch)
    x, _, _ = batch.lm_fit(n_iter=n_iter, lam0=lam0, early_exit=early_exit, graph=graph)
    return batch.split(x)
