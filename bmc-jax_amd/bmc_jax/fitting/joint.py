"""Fits in which samples share values, or a value is free in some samples only
(src/bmc_jax/fitting/joint.py:1-453).

The model and its Jacobian come from the GPU kernel: one launch per block of offsets simulates every
sample (every isochromat of a B0 spread as an extra voxel) with one forward tangent per local entry.
The tangent of a local entry is the kernel's derivative with respect to the slot's packed field,
scaled per sample on the host for "factor" and spread nodes, with the start-state derivative passed
as a seed where m_start is free. The Levenberg-Marquardt steps themselves are float64 on the host:
one joint step for a coupled fit (joint.py:375-416), per-sample damping otherwise (joint.py:359-374).
"""
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


def _layout(spec, n, n_blocks):
    """Where every free value lives (joint.py:70-96): free (n, L), pos (L,), n_loc, G (n, n_loc, P), cols."""
    width = [n_blocks if s.scope == "block" else 1 for s in spec.slots]
    pos = np.concatenate([[0], np.cumsum(width)[:-1]]).astype(int)
    free = np.zeros((n, spec.n), bool)
    cols = []
    for l, s in enumerate(spec.slots):
        who = list(range(n)) if s.samples is None else list(s.samples)
        missing = [i for i in who if not 0 <= i < n]
        if missing:
            raise ValueError(f"{s.name}: samples {missing} do not exist (the fit has {n})")
        free[who, l] = True
        if s.scope == "shared":
            cols.append((l, who, -1))
        elif s.scope == "block":
            cols += [(l, who, b) for b in range(n_blocks)]
        else:
            cols += [(l, [i], -1) for i in who]
    G = np.zeros((n, int(sum(width)), len(cols)))
    for p, (l, who, b) in enumerate(cols):
        G[who, pos[l] + max(b, 0), p] = 1.0
    return free, pos, int(sum(width)), G, cols


def _carried(params, m0, spec, sig_held=None):
    """(n, L): each slot's number where it is not free (joint.py:99-117)."""
    out = []
    n = m0.shape[0]
    for s in spec.slots:
        if s.field == B0_SPREAD:
            out.append(np.asarray(sig_held, np.float64))
        elif s.mode == "factor":
            out.append(np.ones(n))
        elif s.mode == "add":
            out.append(np.zeros(n))
        elif s.field == M_START:
            nc = n_cest_of(params)
            out.append(np.asarray(m0[:, 2 * (1 + nc)], np.float64) / np.asarray(params.water_f, np.float64))
        else:
            a = np.asarray(getattr(params, s.field), np.float64)
            out.append(a if s.index is None else a[:, s.index])
    return np.stack(out, axis=1)


def _cov_to_sd_corr(cov, scale):
    sd = np.sqrt(np.clip(np.diag(cov), 0.0, None))
    with np.errstate(invalid="ignore", divide="ignore"):
        corr = cov / np.outer(sd, sd)
    return sd * scale, corr


def _lm_trial(x, lam, J, r, lo, hi, active_bounds=False):
    """The clipped trial point of one damped Gauss-Newton step (joint.py:149-162), float64."""
    x, J, r = np.asarray(x, np.float64), np.asarray(J, np.float64), np.asarray(r, np.float64)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    g, H = J.T @ r, J.T @ J
    eye = np.eye(x.shape[0])
    if active_bounds:
        keep = np.where(((x <= lo) & (g > 0)) | ((x >= hi) & (g < 0)), 0.0, 1.0)
        H = H * keep[:, None] * keep[None, :] + np.diag(1.0 - keep)
        g = g * keep
    return np.clip(x + np.linalg.solve(H + float(lam) * (np.diag(np.diag(H)) + 1e-9 * eye), -g), lo, hi)


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
        n = self.n
        if tan_slots:
            dm0 = np.stack(seeds, axis=1)
            m, Jb = _engine.simulate_with_jac(q, ev, m0.astype(self.dt), tan_slots, [1.0] * len(tan_slots), self.mz,
                                              dM_init=dm0, n_voxels=n, dtype=self.dt)
        else:
            m = _engine.simulate_states(q, ev, m0.astype(self.dt), n_voxels=n, dtype=self.dt)
            Jb = np.zeros((n, m.shape[1], 0))
        return np.asarray(m[..., self.mz], np.float64), np.asarray(Jb, np.float64)

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
    if np.shape(params.water_r1)[0] != n:
        raise ValueError(f"params carry {np.shape(params.water_r1)[0]} samples, z_meas {n}")
    params = SimParams(*[np.array(f) for f in params])
    m0 = np.asarray(experiment.m0)
    if m0.ndim == 1:
        m0 = np.broadcast_to(m0, (n,) + m0.shape)
    dt = np.asarray(params.water_r1).dtype
    z = z.astype(dt)
    events, mz = experiment.events, experiment.mz_loc
    unknown = [s.name for s in spec.slots if s.field not in PSEUDO and s.field not in SimParams._fields]
    if unknown:
        raise ValueError(f"{unknown}: not a SimParams field nor one of {PSEUDO}")
    l_sp = [l for l, s in enumerate(spec.slots) if s.field == B0_SPREAD]
    if l_sp and spread is None:
        raise ValueError(f"{[spec.slots[l].name for l in l_sp]}: b0_spread needs fit_joint(..., spread=B0Spread(...))")
    if len(l_sp) > 1:
        raise ValueError(f"{[spec.slots[l].name for l in l_sp]}: one b0_spread slot at most")
    l_sp = l_sp[0] if l_sp else None
    sig_held = None if spread is None else np.broadcast_to(np.asarray(spread.sigma_hz, np.float64), (n,))
    if block_jacobian not in ("full", "split"):
        raise ValueError(f"block_jacobian {block_jacobian!r}; use 'full' or 'split'")
    blocked = any(s.scope == "block" for s in spec.slots)
    if blocked and (not n_per_block or n_pts % n_per_block):
        raise ValueError(f"a per-block value needs n_per_block dividing the {n_pts} offsets (got {n_per_block})")
    nb = n_pts // n_per_block if blocked else 1
    free, pos, n_loc, G, cols = _layout(spec, n, nb)
    P = G.shape[2]
    L = spec.n
    scales = spec.scales
    is_block = np.array([s.scope == "block" for s in spec.slots])
    carried = _carried(params, m0, spec, sig_held).astype(dt).astype(np.float64)
    X0 = np.broadcast_to(spec.x0_phys, free.shape) if x0_phys is None else np.asarray(x0_phys, float)
    if X0.shape != free.shape:
        raise ValueError(f"x0_phys has shape {X0.shape}, expected {free.shape} (samples, parameters)")
    X0 = np.where(free, X0, carried)

    model = _Model(params, m0, events, mz, spec, free, carried, spread, l_sp, sig_held, blocked, nb, n_per_block,
                   pos, n_loc)
    locs = lambda x: np.einsum("ijp,p->ij", G, np.asarray(x, np.float64))

    def resid_all(x):
        Zm, _ = model.eval(locs(x), want_jac=False)
        return (Zm - z).reshape(-1)

    def jac_all(x):
        _, Jl = model.eval(locs(x))
        return np.einsum("inj,ijp->inp", Jl, G).reshape(-1, P)

    x_start = np.asarray([X0[who[0], l] / scales[l] for l, who, _ in cols], np.float64)
    if _internals:
        return dict(resid=resid_all, jac=jac_all, x0=x_start, cols=cols, model=model, locs=locs)

    coupled = any(s.scope in ("shared", "block") for s in spec.slots)
    if not coupled and (stop_lam is not None or active_bounds or keep_jacobian):
        raise ValueError("stop_lam, active_bounds and keep_jacobian apply to a coupled fit (a shared or "
                         "per-block value); without one every sample has its own damping, as in `fit`")
    nan_steps = 0
    if not coupled:
        # block diagonal: every sample alone with its own damping -- fit's algorithm (joint.py:359-374)
        hold = None if free.all() else free.astype(np.float64)
        lo = spec.lo_phys / scales
        hi = spec.hi_phys / scales
        x = (X0 / scales).astype(dt).astype(np.float64)
        lam = np.full(n, float(lam0))
        eye = np.eye(L)
        cost_traj = []
        for _ in range(int(n_iter)):
            Zm, Jl = model.eval(x)
            r = Zm - z
            c = np.sum(r * r, axis=1)
            if hold is not None:
                Jl = Jl * hold[:, None, :]
            g = np.einsum("nop,no->np", Jl, r)
            H = np.einsum("nop,noq->npq", Jl, Jl)
            A = H + lam[:, None, None] * (H * eye + 1e-9 * eye)
            if hold is not None:
                A = A + (1.0 - hold)[:, :, None] * eye
            delta = np.linalg.solve(A, -g[..., None])[..., 0]
            xn = np.clip(x + delta, lo, hi)
            Zn, _ = model.eval(xn, want_jac=False)
            cn = np.sum((Zn - z) ** 2, axis=1)
            better = cn < c
            x = np.where(better[:, None], xn, x)
            lam = np.where(better, lam * 0.5, lam * 3.0)
            cost_traj.append(np.where(better, cn, c))
        loc = x
        cost = np.asarray(cost_traj, float).sum(axis=1) if cost_traj else np.zeros(0)
    else:
        # coupled: one joint step over every free value, one damping (joint.py:375-416)
        col_slot = np.array([l for l, _, _ in cols], int)
        col_scale = scales[col_slot]
        lo = spec.lo_phys[col_slot] / col_scale
        hi = spec.hi_phys[col_slot] / col_scale
        x = x_start.astype(dt).astype(np.float64)
        lam, cost = float(lam0), []
        kept = None
        for it in range(int(n_iter)):
            if kept is None:
                r, J = resid_all(x), jac_all(x)
                finite = bool(np.all(np.isfinite(J)))
            else:
                (r, J), finite = kept, True
            if finite:
                c = float(r @ r)
                xn = _lm_trial(x, lam, J, r, lo, hi, active_bounds)
                cn = float(np.sum(resid_all(xn) ** 2))
                better = cn < c
                if better:
                    x, c = xn, cn
                lam = lam * 0.5 if better else lam * 3.0
                cost.append(c)
                kept = (r, J) if keep_jacobian and not better else None
            else:
                nan_steps += 1
                lam = lam * 3.0
                cost.append(float(r @ r))
            if callback is not None:
                callback(it, cost[-1])
            if stop_lam is not None and lam > stop_lam:
                break
        loc = locs(x)
        cost = np.asarray(cost)

    loc_np = np.asarray(loc, float)
    X = np.where(free, loc_np[:, pos] * scales, carried)
    z_fit, Jl = model.eval(loc_np)
    r_all = z_fit - np.asarray(z, float)
    sd = np.full(free.shape, np.nan)
    corr = np.full(free.shape + (L,), np.nan)
    blocks = {}
    if not coupled:
        for i in range(n):
            use = np.flatnonzero(free[i])
            if use.size:
                Ju = Jl[i][:, pos[use]]
                s2 = float(r_all[i] @ r_all[i]) / max(r_all.shape[1] - use.size, 1)
                sd[i, use], corr[i][np.ix_(use, use)] = _cov_to_sd_corr(np.linalg.pinv(Ju.T @ Ju) * s2, scales[use])
    else:
        J = np.einsum("inj,ijp->inp", Jl, G).reshape(-1, P)
        s2 = float(np.sum(r_all ** 2)) / max(r_all.size - P, 1)
        sd_col, corr_col = _cov_to_sd_corr(np.linalg.pinv(J.T @ J) * s2, col_scale)
        for i in range(n):
            ps = [p for p, (l, who, b) in enumerate(cols) if i in who and b < 0]
            ls = col_slot[ps]
            sd[i, ls] = sd_col[ps]
            corr[i][np.ix_(ls, ls)] = corr_col[np.ix_(ps, ps)]
        for l in np.flatnonzero(is_block):
            ps = [p for p, (ll, _, b) in enumerate(cols) if ll == l]
            xb = np.asarray(x, float)[ps] * scales[l]
            blocks[spec.slots[l].name] = dict(x=xb, sd=sd_col[ps], at_bounds=(
                np.isclose(xb, spec.slots[l].lo) | np.isclose(xb, spec.slots[l].hi)))
    X[:, is_block] = np.nan
    X0 = np.where(is_block[None, :], np.nan, X0)
    free = free & ~is_block[None, :]
    return FitResult(X, z_fit, np.asarray(z, float), spec, n_per_block, free=free, sd=sd, corr=corr, x0=X0, cost=cost,
                     nan_steps=nan_steps, blocks=blocks or None)
