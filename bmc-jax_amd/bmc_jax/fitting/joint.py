"""Fits in which samples share values, or a value is free in some samples only
(src/bmc_jax/fitting/joint.py:165-453).

The model and its Jacobian come from the GPU kernel: one launch per block of offsets simulates every
sample (every isochromat of a B0 spread as an extra voxel) with one forward tangent per local entry.
The tangent of a local entry is the kernel's derivative with respect to the slot's packed field,
scaled per sample on the host for "factor" and spread nodes, with the start-state derivative passed
as a seed where m_start is free.

The optimiser follows the reference's split (joint.py:357-416):

* Separable fits (no shared and no per-block value) give every sample its own damping, `fit`'s LM.
  When every slot is a plain value of a SimParams field and there is no B0 spread, the whole fit is
  one `bmc_lm_fit` call with the free mask (joint.py:359-374 vmaps levenberg_marquardt the same way).
  Factors, offsets, the start state and spreads have no kernel-side form of their write, so those
  separable fits step on the host in float64 around GPU model evaluations.
* Coupled fits take one damped step over the joint vector (joint.py:375-416): GPU model and
  Jacobian, float64 step on the host.
"""
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .. import _engine, _native
from ..simulation.events import slice_rows
from ..simulation.params import SimParams, initial_state, n_cest_of
from ..simulation.voxels import tile_params
from .result import FitResult

M_START = "m_start"
B0_SPREAD = "b0_spread"
PSEUDO = (M_START, B0_SPREAD)
_FRACTION_FIELDS = {"water_f", "cest_f", "mt_f"}


@dataclass
class _Unknowns:
    """The joint vector of a fit and how it feeds every sample's local vector (joint.py:70-96).

    A sample's local vector has one entry per slot, or one per block for a per-block slot, starting at
    `first[l]`. Joint column p is `columns[p] = (slot, samples it is free in, block or -1)`; `feed[i, j, p]`
    is 1 where column p sets local entry j of sample i."""
    free: np.ndarray
    first: np.ndarray
    n_local: int
    columns: List[Tuple[int, List[int], int]]
    feed: np.ndarray

    @property
    def n_joint(self):
        return len(self.columns)

    def local(self, x):
        """Joint vector (P,) -> local vectors (n, n_local)."""
        return np.einsum("ijp,p->ij", self.feed, np.asarray(x, np.float64))

    def joint_jacobian(self, Jl):
        """Local Jacobians (n, points, n_local) -> the joint one (n * points, P)."""
        return np.einsum("inj,ijp->inp", Jl, self.feed).reshape(-1, self.n_joint)


def _unknowns(spec, n, n_blocks) -> _Unknowns:
    per_slot = np.array([n_blocks if s.scope == "block" else 1 for s in spec.slots], int)
    first = np.cumsum(per_slot) - per_slot
    free = np.zeros((n, spec.n), bool)
    columns = []
    for l, s in enumerate(spec.slots):
        members = list(range(n)) if s.samples is None else [int(i) for i in s.samples]
        outside = [i for i in members if i < 0 or i >= n]
        if outside:
            raise ValueError(f"{s.name}: samples {outside} do not exist in a fit of {n} samples")
        free[members, l] = True
        if s.scope == "sample":
            columns.extend((l, [i], -1) for i in members)
        elif s.scope == "shared":
            columns.append((l, members, -1))
        else:
            columns.extend((l, members, b) for b in range(n_blocks))
    feed = np.zeros((n, int(per_slot.sum()), len(columns)))
    for p, (l, members, b) in enumerate(columns):
        feed[members, first[l] + max(b, 0), p] = 1.0
    return _Unknowns(free, first, int(per_slot.sum()), columns, feed)


def _held_values(params, m0, spec, held_sigma=None):
    """(n, L): the number each slot has where it is not fitted (joint.py:99-117): the field's own value,
    the neutral factor 1, the neutral offset 0, for m_start the start state's water Mz relative to the
    water fraction, for b0_spread the held spread width."""
    n = m0.shape[0]
    table = np.empty((n, spec.n))
    water_mz = 2 * (1 + n_cest_of(params))
    for l, s in enumerate(spec.slots):
        if s.field == B0_SPREAD:
            table[:, l] = np.asarray(held_sigma, np.float64)
        elif s.mode in ("factor", "add"):
            table[:, l] = 1.0 if s.mode == "factor" else 0.0
        elif s.field == M_START:
            table[:, l] = np.asarray(m0[:, water_mz], np.float64) / np.asarray(params.water_f, np.float64)
        else:
            a = np.asarray(getattr(params, s.field), np.float64)
            table[:, l] = a if s.index is None else a[:, s.index]
    return table


def _errors(cov, scale):
    """Covariance in scaled units -> (standard deviations in physical units, correlation matrix)."""
    sd = np.sqrt(np.maximum(np.diag(cov), 0.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        corr = cov / sd[:, None] / sd[None, :]
    return sd * scale, corr


def _lm_trial(x, lam, J, r, lo, hi, active_bounds=False):
    """The clipped trial point of one damped Gauss-Newton step (joint.py:149-162), float64.

    active_bounds: a value on a bound whose step leaves the box is held for this step (identity row and
    column, zero gradient)."""
    x, J, r = np.asarray(x, np.float64), np.asarray(J, np.float64), np.asarray(r, np.float64)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    g = J.T @ r
    H = J.T @ J
    if active_bounds:
        railed = ((x <= lo) & (g > 0)) | ((x >= hi) & (g < 0))
        H[railed, :] = 0.0
        H[:, railed] = 0.0
        H[railed, railed] = 1.0
        g = np.where(railed, 0.0, g)
    A = H + float(lam) * np.diag(np.diag(H) + 1e-9)
    return np.clip(x + np.linalg.solve(A, -g), lo, hi)


class _Model:
    """Z and dZ/d(local entries) for every sample, on the GPU."""

    def __init__(self, params, m0, events, mz, spec, free, carried, spread, l_sp, sig_held, blocked, nb, n_per_block,
                 pos, n_loc):
        self.params, self.m0, self.events, self.mz, self.spec = params, m0, events, mz, spec
        self.free, self.carried, self.spread, self.l_sp, self.sig_held = free, carried, spread, l_sp, sig_held
        self.blocked, self.nb, self.npb, self.pos, self.n_loc = blocked, nb, n_per_block, pos, n_loc
        self.n = m0.shape[0]
        self.nc = n_cest_of(params)
        self.dt = np.asarray(params.water_r1).dtype
        self.L = spec.n
        self.scales = spec.scales
        self.n_pts = np.shape(events.dt)[0]
        self.ev_blocks = ([slice_rows(events, slice(b * n_per_block, (b + 1) * n_per_block)) for b in range(nb)]
                          if blocked else [events])

    def _index(self, l, b):
        s = self.spec.slots[l]
        return self.pos[l] + (b if s.scope == "block" else 0)

    def _apply(self, eff):
        """Per-sample params, M0 and spread width with every slot's number applied (joint.py:120-138).

        eff (n, L) physical. Returns (params, m0 (n, N), base values per slot (n, L), m_rebuild (n, N) or None)."""
        p = self.params
        upd = {}
        base = np.zeros((self.n, self.L))
        m_slot = None
        for l, s in enumerate(self.spec.slots):
            if s.field == M_START:
                m_slot = l
                continue
            if s.field == B0_SPREAD:
                continue
            cur = np.array(upd.get(s.field, np.asarray(getattr(p, s.field))), dtype=self.dt, copy=True)
            old = cur if s.index is None else cur[:, s.index]
            base[:, l] = np.asarray(old, np.float64)
            new = eff[:, l] if s.mode == "value" else (old * eff[:, l] if s.mode == "factor" else old + eff[:, l])
            if s.index is None:
                cur = np.asarray(new, self.dt)
            else:
                cur[:, s.index] = new
            upd[s.field] = cur
        q = p._replace(**upd)
        m0 = np.asarray(self.m0, np.float64)
        eq = None
        if m_slot is not None:
            m = eff[:, m_slot]
            if self.spec.slots[m_slot].mode == "value":
                eq = np.asarray(initial_state(q), np.float64)
                rebuilt = eq * m[:, None]
            else:
                rebuilt = m0 * m[:, None]
            m0 = np.where(self.free[:, m_slot][:, None], rebuilt, m0)
        return q, m0, base, m_slot, eq

    def eval(self, loc, want_jac=True):
        """loc (n, n_loc) scaled -> Z (n, n_pts) and J (n, n_pts, n_loc)."""
        n, L, nc = self.n, self.L, self.nc
        loc = np.asarray(loc, np.float64)
        N = 3 * (1 + nc) + 1
        Z = np.zeros((n, self.n_pts))
        J = np.zeros((n, self.n_pts, self.n_loc)) if want_jac else None
        for b, ev in enumerate(self.ev_blocks):
            idx = [self._index(l, b) for l in range(L)]
            cs = self.carried
            eff = np.where(self.free, (loc[:, idx] * self.scales).astype(self.dt), cs.astype(self.dt)).astype(np.float64)
            q, m0, base, m_slot, eq = self._apply(eff)
            # tangents: one per slot that is free in some sample (its derivative w.r.t. the slot's value)
            tan_slots, tan_l, seeds = [], [], []
            for l, s in enumerate(self.spec.slots):
                if not self.free[:, l].any() or not want_jac:
                    continue
                seed = np.zeros((n, N))
                if s.field == M_START:
                    tan_slots.append(-1)
                    seed += (eq if s.mode == "value" else np.asarray(self.m0, np.float64)) * self.free[:, l][:, None]
                elif s.field == B0_SPREAD:
                    tan_slots.append(_native.slot_index("b0_inhom", None, nc))
                else:
                    tan_slots.append(_native.slot_index(s.field, s.index, nc))
                    if m_slot is not None and self.spec.slots[m_slot].mode == "value" and s.field in _FRACTION_FIELDS:
                        # initial_state(q) * m moves with the pool fraction it places (params.py:197-212)
                        mfac = eff[:, m_slot] * self.free[:, m_slot]
                        zi = (2 * (1 + nc) if s.field == "water_f" else
                              3 * (1 + nc) if s.field == "mt_f" else 2 * (1 + nc) + 1 + int(s.index))
                        seed[:, zi] += mfac
                tan_l.append(l)
                seeds.append(seed)
            if self.spread is None:
                z, Jb = self._run(q, ev, m0, tan_slots, seeds)
            else:
                z, Jb = self._run_spread(q, ev, m0, eff, tan_slots, seeds, tan_l)
            rows = slice(b * self.npb, (b + 1) * self.npb) if self.blocked else slice(None)
            Z[:, rows] = z
            if want_jac:
                for t, l in enumerate(tan_l):
                    s = self.spec.slots[l]
                    # chain to the scaled local entry: d value / d loc = scale (x base for a factor)
                    mult = self.scales[l] * self.free[:, l].astype(np.float64)
                    if s.mode == "factor" and s.field != M_START:
                        mult = mult * base[:, l]
                    J[:, rows, self._index(l, b)] = Jb[:, :, t] * mult[:, None]
        return Z, J

    def _run(self, q, ev, m0, tan_slots, seeds):
        return self._run_inner(q, ev, m0, tan_slots, seeds, self.n)

    def _run_spread(self, q, ev, m0, eff, tan_slots, seeds, tan_l):
        """Every sample's isochromats as extra voxels; weighted mean (spread.py:77-106)."""
        n = self.n
        u, w = self.spread.nodes_weights()
        K = u.size
        sig = eff[:, self.l_sp] if self.l_sp is not None else np.asarray(self.sig_held, np.float64)
        hz = np.asarray(q.gamma, np.float64) * np.asarray(q.b0, np.float64) / (2 * np.pi)
        hz = np.broadcast_to(hz, (n,))
        rep = lambda a: np.repeat(np.asarray(a), K, axis=0)
        fields = {f: rep(getattr(q, f)) for f in SimParams._fields}
        shift = (sig[:, None] * u[None, :] / hz[:, None]).reshape(-1)
        fields["b0_inhom"] = (np.asarray(fields["b0_inhom"], np.float64) + shift).astype(self.dt)
        qq = SimParams(**fields)
        m0r = rep(m0)
        seeds_r = [rep(s) for s in seeds]
        z, Jb = self._run_inner(qq, ev, m0r, tan_slots, seeds_r, n * K)
        z = np.einsum("nk,nko->no", np.broadcast_to(w, (n, K)), z.reshape(n, K, -1))
        if Jb.shape[-1]:
            Jb = Jb.reshape(n, K, Jb.shape[1], -1)
            out = np.zeros((n, Jb.shape[2], Jb.shape[3]))
            for t, l in enumerate(tan_l):
                wk = w[None, :] if self.spec.slots[l].field != B0_SPREAD else w[None, :] * u[None, :] / hz[:, None]
                out[:, :, t] = np.einsum("nk,nko->no", np.broadcast_to(wk, (n, K)), Jb[:, :, :, t])
            Jb = out
        else:
            Jb = np.zeros((n, z.shape[1], 0))
        return z, Jb

    def _run_inner(self, q, ev, m0, tan_slots, seeds, V):
        if tan_slots:
            m, Jb = _engine.simulate_with_jac(q, ev, m0.astype(self.dt), tan_slots, [1.0] * len(tan_slots), self.mz,
                                              dM_init=np.stack(seeds, axis=1), n_voxels=V, dtype=self.dt)
        else:
            m = _engine.simulate_states(q, ev, m0.astype(self.dt), n_voxels=V, dtype=self.dt)
            Jb = np.zeros((V, m.shape[1], 0))
        return np.asarray(m[..., self.mz], np.float64), np.asarray(Jb, np.float64)


def _kernel_writes_every_slot(spec, spread):
    """True when bmc_lm_fit can run the separable fit itself: its kernels write a scaled value into a
    packed field, which is exactly a "value" slot of a SimParams field with no spread around it."""
    return spread is None and all(s.mode == "value" and s.field not in PSEUDO for s in spec.slots)


def _separable_on_device(params, m0, events, mz, spec, z, x0, free, n_iter, lam0, dt):
    """Every sample's own LM in one bmc_lm_fit call (joint.py:359-374). Returns (x (n, L), summed cost)."""
    from .independent import _lm_batched
    mask = None if free.all() else free.astype(np.float64)
    x, ctraj, _, _ = _lm_batched(params, events, m0, mz, spec, z, x0, int(n_iter), dt, lam0=lam0, free=mask,
                                 want_cost=True)
    cost = ctraj.sum(axis=0) if ctraj is not None else np.zeros(0)
    return x, cost


def _separable_on_host(model, z, x, lo, hi, free, n_iter, lam0):
    """Every sample's own LM (lm.py:43-61 per sample), float64 steps around GPU model evaluations."""
    mask = None if free.all() else free.astype(np.float64)
    n, L = x.shape
    damp = np.full(n, float(lam0))
    eye = np.eye(L)
    history = []
    for _ in range(int(n_iter)):
        Zm, Jl = model.eval(x)
        r = Zm - z
        c = np.einsum("no,no->n", r, r)
        if mask is not None:
            Jl = Jl * mask[:, None, :]
        g = np.einsum("nop,no->np", Jl, r)
        H = np.einsum("nop,noq->npq", Jl, Jl)
        A = H + damp[:, None, None] * (H * eye + 1e-9 * eye)
        if mask is not None:
            A = A + (1.0 - mask)[:, :, None] * eye
        trial = np.clip(x + np.linalg.solve(A, -g[..., None])[..., 0], lo, hi)
        Zt, _ = model.eval(trial, want_jac=False)
        ct = np.einsum("no,no->n", Zt - z, Zt - z)
        accept = ct < c
        x = np.where(accept[:, None], trial, x)
        damp = damp * np.where(accept, 0.5, 3.0)
        history.append(np.where(accept, ct, c))
    return x, (np.asarray(history).sum(axis=1) if history else np.zeros(0))


def _coupled(resid, jac, x, lo, hi, n_iter, lam0, active_bounds, keep_jacobian, stop_lam, callback):
    """One damping over the joint vector (joint.py:375-416). Returns (x, cost after every step, NaN steps)."""
    lam = float(lam0)
    costs, n_nan = [], 0
    reuse = None                     # (r, J) of the current point after a rejected step (keep_jacobian)
    for step in range(int(n_iter)):
        r, J = reuse if reuse is not None else (resid(x), jac(x))
        reuse = None
        c = float(r @ r)
        if np.isfinite(J).all():
            trial = _lm_trial(x, lam, J, r, lo, hi, active_bounds)
            c_trial = float(np.sum(resid(trial) ** 2))
            if c_trial < c:
                x, c, lam = trial, c_trial, lam * 0.5
            else:
                lam *= 3.0
                if keep_jacobian:
                    reuse = (r, J)
        else:                        # a Jacobian with NaN or inf: no step, harder damping, counted
            n_nan += 1
            lam *= 3.0
        costs.append(c)
        if callback is not None:
            callback(step, c)
        if stop_lam is not None and lam > stop_lam:
            break
    return x, np.asarray(costs), n_nan


def fit_joint(experiment, spec, z_meas, n_iter=30, x0_phys=None, n_per_block=None, max_chunk=None, lam0=1e-2,
              callback=None, spread=None, block_jacobian="full", stop_lam=None, active_bounds=False,
              keep_jacobian=False, _internals=False):
    """Fit every sample at once: shared, per-block and partly free values allowed (joint.py:165-453).

    max_chunk and block_jacobian are accepted for signature compatibility: the kernel computes every
    local tangent exactly in one launch per block, so both choices give the same Jacobian.
    """
    z = np.atleast_2d(np.asarray(z_meas))
    n, n_pts = z.shape
    params = experiment.params
    if np.ndim(params.water_r1) == 0:
        params = tile_params(params, n)
    n_given = np.shape(params.water_r1)[0]
    if n_given != n:
        raise ValueError(f"z_meas has {n} samples but the params describe {n_given}")
    params = SimParams(*[np.array(f) for f in params])
    m0 = np.asarray(experiment.m0)
    if m0.ndim == 1:
        m0 = np.broadcast_to(m0, (n,) + m0.shape)
    dt = np.asarray(params.water_r1).dtype
    z = z.astype(dt)
    events, mz = experiment.events, experiment.mz_loc

    unknown = [s.name for s in spec.slots if s.field not in PSEUDO and s.field not in SimParams._fields]
    if unknown:
        raise ValueError(f"{unknown} name neither a SimParams field nor one of {PSEUDO}")
    spread_slots = [l for l, s in enumerate(spec.slots) if s.field == B0_SPREAD]
    names = [spec.slots[l].name for l in spread_slots]
    if spread_slots and spread is None:
        raise ValueError(f"{names}: a b0_spread value needs fit_joint(..., spread=B0Spread(...)) to act on")
    if len(spread_slots) > 1:
        raise ValueError(f"{names}: at most one b0_spread value per fit")
    l_sp = spread_slots[0] if spread_slots else None
    sig_held = None if spread is None else np.broadcast_to(np.asarray(spread.sigma_hz, np.float64), (n,))
    if block_jacobian not in ("full", "split"):
        raise ValueError(f"block_jacobian must be 'full' or 'split', not {block_jacobian!r}")
    blocked = any(s.scope == "block" for s in spec.slots)
    if blocked and (not n_per_block or n_pts % n_per_block):
        raise ValueError(f"per-block values need n_per_block to divide the {n_pts} offsets; got {n_per_block}")
    n_blocks = n_pts // n_per_block if blocked else 1

    unk = _unknowns(spec, n, n_blocks)
    free = unk.free
    P, L = unk.n_joint, spec.n
    scales = spec.scales
    is_block = np.array([s.scope == "block" for s in spec.slots])
    carried = _held_values(params, m0, spec, sig_held).astype(dt).astype(np.float64)
    X0 = np.broadcast_to(spec.x0_phys, free.shape) if x0_phys is None else np.asarray(x0_phys, float)
    if X0.shape != free.shape:
        raise ValueError(f"x0_phys must have shape {free.shape} (samples, parameters), not {X0.shape}")
    X0 = np.where(free, X0, carried)

    model = _Model(params, m0, events, mz, spec, free, carried, spread, l_sp, sig_held, blocked, n_blocks,
                   n_per_block, unk.first, unk.n_local)

    def resid_all(x):
        Zm, _ = model.eval(unk.local(x), want_jac=False)
        return (Zm - z).reshape(-1)

    def jac_all(x):
        _, Jl = model.eval(unk.local(x))
        return unk.joint_jacobian(Jl)

    # a shared or per-block value starts at its first member's entry
    x_start = np.array([X0[members[0], l] / scales[l] for l, members, _ in unk.columns], np.float64)
    if _internals:
        return dict(resid=resid_all, jac=jac_all, x0=x_start, cols=unk.columns, model=model, locs=unk.local)

    coupled = any(s.scope != "sample" for s in spec.slots)
    if not coupled and (stop_lam is not None or active_bounds or keep_jacobian):
        raise ValueError("stop_lam, active_bounds and keep_jacobian only act on a coupled fit (one with a "
                         "shared or per-block value); a separable fit damps every sample alone, as fit does")
    nan_steps = 0
    if not coupled:
        x0 = (X0 / scales).astype(dt).astype(np.float64)
        if _kernel_writes_every_slot(spec, spread):
            loc, cost = _separable_on_device(params, m0, events, mz, spec, z, x0, free, n_iter, lam0, dt)
        else:
            loc, cost = _separable_on_host(model, z, x0, spec.lo_phys / scales, spec.hi_phys / scales, free,
                                           n_iter, lam0)
        col_slot = col_scale = x = None
    else:
        col_slot = np.array([l for l, _, _ in unk.columns], int)
        col_scale = scales[col_slot]
        x, cost, nan_steps = _coupled(resid_all, jac_all, x_start.astype(dt).astype(np.float64),
                                      spec.lo_phys[col_slot] / col_scale, spec.hi_phys[col_slot] / col_scale,
                                      n_iter, lam0, active_bounds, keep_jacobian, stop_lam, callback)
        loc = unk.local(x)

    loc = np.asarray(loc, float)
    X = np.where(free, loc[:, unk.first] * scales, carried)
    z_fit, Jl = model.eval(loc)
    misfit = z_fit - np.asarray(z, float)
    sd = np.full(free.shape, np.nan)
    corr = np.full(free.shape + (L,), np.nan)
    blocks = {}
    if not coupled:
        # misfit-scaled error bars per sample: s^2 = cost_i / (points - its free values)
        for i in np.flatnonzero(free.any(axis=1)):
            mine = np.flatnonzero(free[i])
            Ji = Jl[i][:, unk.first[mine]]
            s2 = float(misfit[i] @ misfit[i]) / max(n_pts - mine.size, 1)
            sd[i, mine], corr[i][np.ix_(mine, mine)] = _errors(np.linalg.pinv(Ji.T @ Ji) * s2, scales[mine])
    else:
        # one covariance over the joint vector, s^2 = cost / (all points - joint values)
        Jj = unk.joint_jacobian(Jl)
        s2 = float(np.sum(misfit ** 2)) / max(misfit.size - P, 1)
        sd_j, corr_j = _errors(np.linalg.pinv(Jj.T @ Jj) * s2, col_scale)
        for i in range(n):
            own = [p for p, (_, members, b) in enumerate(unk.columns) if b < 0 and i in members]
            slots_i = col_slot[own]
            sd[i, slots_i] = sd_j[own]
            corr[i][np.ix_(slots_i, slots_i)] = corr_j[np.ix_(own, own)]
        for l in np.flatnonzero(is_block):
            s = spec.slots[l]
            own = [p for p, col in enumerate(unk.columns) if col[0] == l]
            xb = np.asarray(x, float)[own] * scales[l]
            blocks[s.name] = dict(x=xb, sd=sd_j[own], at_bounds=np.isclose(xb, s.lo) | np.isclose(xb, s.hi))
    # per-block values are reported in `blocks`; their column of x is NaN
    X[:, is_block] = np.nan
    X0 = np.where(is_block[None, :], np.nan, X0)
    return FitResult(X, z_fit, np.asarray(z, float), spec, n_per_block, free=free & ~is_block[None, :], sd=sd,
                     corr=corr, x0=X0, cost=cost, nan_steps=nan_steps, blocks=blocks or None)
