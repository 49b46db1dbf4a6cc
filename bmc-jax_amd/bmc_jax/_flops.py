"""Algorithmic FLOP counts of the simulate kernel (csrc/bmc_sim.hip), for roofline reporting.

`executed_flops` replays, on the host, the kernel's per-event choice of Taylor degree m and squaring
count s (the same 1-norm and theta_m table) and counts the useful float64 operations of every step:
  generator apply X v       29 NC + 27 flops (13 NC + 13 FMA, 3 NC + 1 multiplies)
  tangent apply dX v        per slot kind (below), plus X dv
  Taylor term bookkeeping   2 N per vector (scale by 1/k, accumulate)
  dense product (column)    2 N^2 (1 + 2 PT) per column, N + 1 useful columns (squarings, and chaining
                            an event that needed squarings; other chained events need none)
  block product (column)    2 (NXY^2 + NZ^2) (1 + 2 PT) per column for an event without RF (rf_amp 0):
                            its propagator is block-diagonal in the transverse rows (NXY = 2 (1 + NC))
                            and the z rows (NZ = N - NXY), and the kernel multiplies the blocks apart
  chained block applied     count times for count < 4; from 4 on by binary powering: popcount(count)
                            applications and bit_length(count) - 1 squarings of the chained propagator
  apply to the state        (N + 1)(N + 3 PT N) + N (1 + PT) N adds of the team reduction
  palindrome factorisation  a chained run opening with 2h mirrored events (plan level 3; teams with a
                            spare lane, plans without kept pulses) chains h events on N + 2 columns
                            (the spare lane's reversed chain included) and recombines them with
                            2 N^2 (1 + 2 PT) flops per column, when the voxel qualifies (mirror_ok in
                            csrc/bmc_device.h) and no event of the half needs squarings
Only the N + 1 columns of a propagator count, not the idle lanes of a 16-lane team.
`reference_flops` is a model of the reference's per-event work (every event its own
Pade exponential with its tangents, reference float32 degree/squaring rule).
"""
import math

import numpy as np

THETA = np.array([0.0, 2.22e-16, 2.581e-08, 1.386e-05, 0.0003397, 0.002401, 0.009066, 0.02384, 0.04991, 0.08958,
                  0.1442, 0.2142, 0.2996, 0.3998, 0.5139, 0.6411, 0.7803, 0.9305, 1.091, 1.26, 1.438, 1.624, 1.816,
                  2.015, 2.219, 2.429, 2.643, 2.861, 3.084, 3.31, 3.54])
M_MAX = 30

# tangent kinds by packed slot (csrc/bmc_sim.hip tan_kind_of_slot)
_DX_FLOPS_SCALAR = {0: 4, 1: 4, 2: 0, 3: 8, 4: 6, 5: 2, 6: 4, 7: 2, 8: 0}


def dx_flops(slot, nc):
    if slot < 0:
        return 0
    if slot in _DX_FLOPS_SCALAR:
        return _DX_FLOPS_SCALAR[slot]
    if slot in (9, 10):                # b0, gamma
        return 2 * (3 + 3 * nc)
    if slot == 11:                     # b0_inhom
        return 2 * (3 + 2 * nc)
    if slot == 12:                     # rel_b1
        return 2 * (5 + 4 * nc)
    kind = (slot - 13) // nc           # cest_k, cest_f, cest_dw, cest_r1, cest_r2
    return (24, 14, 4, 4, 4)[kind]


def _gen_l1(prm, nc, dt, amp, cph, sph, fr):
    """The kernel's augmented 1-norm per (voxel, event) (build_gen), vectorised. prm [V, S]; events 1-D."""
    r1w, r2w, kmt, fmt, dwmt, r1mt, r2mt = (prm[:, i][:, None] for i in (0, 1, 3, 4, 5, 6, 7))
    ls = prm[:, 8][:, None]
    b0, gam, b0i, relb1 = (prm[:, i][:, None] for i in (9, 10, 11, 12))
    k = prm[:, 13:13 + nc][:, :, None]
    f = prm[:, 13 + nc:13 + 2 * nc][:, :, None]
    dw = prm[:, 13 + 2 * nc:13 + 3 * nc][:, :, None]
    r1c = prm[:, 13 + 3 * nc:13 + 4 * nc][:, :, None]
    r2c = prm[:, 13 + 4 * nc:13 + 5 * nc][:, :, None]
    L = gam * 1e6 * b0
    ks = (prm[:, 13:13 + nc] * prm[:, 13 + nc:13 + 2 * nc]).sum(1)[:, None]
    kwmt = kmt * fmt
    w1 = 2 * np.pi * amp[None, :] * relb1
    w1x, w1y = w1 * cph[None, :], w1 * sph[None, :]
    eff = 2 * np.pi * fr[None, :] - L * b0i * 1e-6
    # the MT saturation term only matters through |dmt| in the norm; w_mt >= 0 adds to it
    from .simulation.matrix import lorentzian, super_lorentzian
    wmt = np.zeros_like(eff)
    on = (fmt[:, 0] > 1e-9) & (ls[:, 0] != 0)
    if on.any():
        for v in np.flatnonzero(on):
            off = eff[v] - dwmt[v, 0] * 1e-6 * L[v, 0]
            t2 = 1.0 / r2mt[v, 0]
            js = np.flatnonzero(w1[v] != 0)
            uniq, inv = np.unique(off[js], return_inverse=True)
            g = np.array([lorentzian(u, t2) if ls[v, 0] == 1 else super_lorentzian(u, t2, L[v, 0] * 1e-6)
                          for u in uniq])
            wmt[v, js] = w1[v, js] ** 2 * g[inv]
    axx, azz, dmt = -r2w - ks, -r1w - ks - kwmt, -r1mt - kmt - wmt
    skw = np.abs(prm[:, 13:13 + nc] * prm[:, 13 + nc:13 + 2 * nc]).sum(1)[:, None]
    n1 = np.maximum.reduce([np.abs(axx) + np.abs(eff) + np.abs(w1y) + skw, np.abs(eff) + np.abs(axx) + np.abs(w1x) + skw,
                            np.abs(w1y) + np.abs(w1x) + np.abs(azz) + skw + np.abs(kwmt), np.abs(kmt) + np.abs(dmt),
                            (np.abs(r1w) + np.abs(r1mt * fmt) + np.abs(r1c * f).sum(1)) * np.ones_like(eff)])
    if nc:
        oi = dw * 1e-6 * L[:, :, None] - eff[:, None, :]
        bi, ci = -r2c - k, -r1c - k
        p1 = (np.abs(k) + np.abs(bi) + np.abs(oi) + np.abs(w1y[:, None, :])).max(1)
        p2 = (np.abs(k) + np.abs(oi) + np.abs(bi) + np.abs(w1x[:, None, :])).max(1)
        p3 = (np.abs(k) + np.abs(w1y[:, None, :]) + np.abs(w1x[:, None, :]) + np.abs(ci)).max(1)
        n1 = np.maximum.reduce([n1, p1, p2, p3])
    return n1 * dt[None, :]


def degree_and_squarings(l1, mmax=M_MAX):
    """The kernels' Taylor degree m and squaring count s for 1-norms l1, with degree cap mmax."""
    l1 = np.asarray(l1, np.float64)
    m = np.searchsorted(THETA[1:mmax + 1], l1, side="left") + 1     # smallest m with theta_m >= l1
    m = np.clip(m, 1, mmax)
    s = np.where(l1 <= THETA[mmax], 0, np.ceil(np.log2(np.maximum(l1, 1e-300) / THETA[mmax])))
    s = np.maximum(s, 0).astype(np.int64)
    big = s > 0
    while True:
        bad = big & (l1 * np.ldexp(1.0, -s) > THETA[mmax])
        if not bad.any():
            break
        s = s + bad
    m = np.where(big, mmax, m)
    return m.astype(np.int64), s


def team_of(nc):
    """Lanes per team of the simulate kernel (one per propagator column, csrc/bmc_sim.hip)."""
    na = 3 * (1 + nc) + 2
    return 8 if na <= 8 else 12 if na <= 12 else 16 if na <= 16 else 32


def mirror_ok(prm, nc, slots):
    """Per voxel of prm [V, S]: whether the parameters and the pass's tangent slots admit the palindrome
    factorisation (csrc/bmc_device.h mirror_ok): every pool and MT has a positive fraction or is
    decoupled (fraction and rate 0) with no tangent along its rate."""
    fmt, kmt = prm[:, 4], prm[:, 3]
    ok = (fmt > 0) | ((fmt == 0) & (kmt == 0))
    if 3 in slots:                                     # mt_k
        ok &= fmt > 0
    for i in range(nc):
        f, k = prm[:, 13 + nc + i], prm[:, 13 + i]
        ok &= (f > 0) | ((f == 0) & (k == 0))
        if 13 + i in slots:                            # cest_k of pool i
            ok &= f > 0
    return ok


def executed_flops(prm, nc, events, ops, row_ptr, slots, pt=1, mirror=True):
    """Useful float64 flops of one bmc_simulate call, per voxel of prm [V, S] (summed over offsets and passes).

    events: dict/SimEvent of (O, T) arrays; ops/row_ptr: the plan's program (kinds 0 SEQ, 1 BLOCK,
    2 KEEP counted as a BLOCK, 3 REUSE: the kept propagator applied once, rotated; column 4 the
    palindromic half); slots: tangent slots. mirror=False for kernels that chain every event (the
    uneven batch).
    """
    ev = {k: np.asarray(getattr(events, k) if not isinstance(events, dict) else events[k])
          for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    N = 3 * (1 + nc) + 1
    NA = N + 1
    NXY = 2 * (1 + nc)
    FX = 29 * nc + 27
    P = len(slots)
    passes = [slots[i:i + pt] for i in range(0, P, pt)] if P else [[]]
    V = prm.shape[0]
    total = np.zeros(V)
    O = ev["dt"].shape[0]
    ops = np.asarray(ops)
    can_mirror = (mirror and ops.shape[1] > 4 and team_of(nc) > NA
                  and not np.any(ops[:, 0] == 2))     # kernels with kept pulses have no stage
    for o in range(O):
        for oi in range(row_ptr[o], row_ptr[o + 1]):
            kind, e0, ln, cnt = (int(v) for v in ops[oi][:4])
            half = int(ops[oi][4]) if ops.shape[1] > 4 else 0
            if kind == 3:                 # REUSE: the kept propagator applied once, rotated (2 x 6 flops per pool)
                for ps in passes:
                    T = len(ps)
                    total += NA * (N + 3 * T * N) + N * (1 + T) * (NA - 1) + 2 * 6 * (1 + nc) * (1 + T)
                continue
            sel = np.arange(e0, e0 + ln)
            dt = ev["dt"][o, sel].astype(np.float64)
            live = dt > 1e-9
            if not live.any():
                continue
            sel = sel[live]
            l1 = _gen_l1(prm, nc, dt[live], ev["rf_amp"][o, sel].astype(np.float64),
                         np.cos(ev["rf_phase"][o, sel].astype(np.float64)),
                         np.sin(ev["rf_phase"][o, sel].astype(np.float64)), ev["rf_freq"][o, sel].astype(np.float64))
            m, s = degree_and_squarings(l1)
            rf0 = ev["rf_amp"][o, sel] == 0          # events without RF: block-diagonal products
            for ps in passes:
                T = len(ps)
                fdx = sum(dx_flops(sl, nc) for sl in ps)
                term = FX + T * FX + fdx + 2 * N * (1 + T)
                mul = NA * 2 * N * N * (1 + 2 * T)
                mulx = np.where(rf0, NA * 2 * (NXY * NXY + (N - NXY) ** 2) * (1 + 2 * T), mul)
                app = NA * (N + 3 * T * N) + N * (1 + T) * (NA - 1)
                if kind == 0:
                    vec = (s == 0)
                    f = np.where(vec, m * term, NA * m * term + s * mulx + app)
                    total += f.sum(axis=1)
                else:
                    # chained: an event without squarings runs its series on the chained columns
                    # directly; one with squarings forms its propagator and multiplies it in
                    f = np.where(s == 0, NA * m * term, NA * m * term + s * mulx + mulx)
                    plain = f.sum(axis=1)
                    if can_mirror and half > 0:
                        # events [e0, e0 + 2h) with dt > 1e-9 are the first live ones of sel
                        nh = int(np.sum(live[:half]))
                        n2 = int(np.sum(live[:2 * half]))
                        ok = mirror_ok(prm, nc, ps) & np.all(s[:, :nh] == 0, axis=1)
                        mir = (NA + 1) * (m[:, :nh] * term).sum(axis=1) + NA * 2 * N * N * (1 + 2 * T) \
                            + f[:, n2:].sum(axis=1)
                        plain = np.where(ok, mir, plain)
                    total += plain + power_flops(cnt, app, mul)
    return total


def power_flops(count, app, mul):
    """Flops of applying a chained propagator count times to the state (csrc/bmc_device.h run_row):
    count applications below 4, else binary powering (popcount applications, bit_length - 1 squarings)."""
    if count < 4:
        return count * app
    return bin(count).count("1") * app + (count.bit_length() - 1) * mul


def reference_flops(n_aug, P, m_ref, s_ref):
    """A model of the reference's per-event work: expm + P tangents + apply (flops)."""
    km = {3: 2, 5: 3, 7: 4, 9: 5, 13: 6}[m_ref]
    n3, n2 = n_aug ** 3, n_aug ** 2
    return 2 * n3 * (km + s_ref) + 8.0 / 3.0 * n3 + P * (4 * n3 * (km + s_ref) + 4 * n3) + 2 * n2 + 4 * P * n2


def reference_equivalent_flops(prm, nc, events, P):
    """The reference's work for the same rows, per voxel of prm [V, S]: every event with dt > 1e-9 its own
    Pade exponential with P tangents and one apply (reference_flops), degree and squarings by the
    reference float32 rule from the event's 1-norm. The "effective" figure beside the executed one."""
    ev = {k: np.asarray(getattr(events, k) if not isinstance(events, dict) else events[k])
          for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    n_aug = 3 * (1 + nc) + 2
    total = np.zeros(prm.shape[0])
    lo, hi, th = 0.4258730016922831, 1.880152677804762, 3.925724783138660
    cost = {m: np.array([reference_flops(n_aug, P, m, s) for s in range(64)]) for m in (3, 5, 7)}
    for o in range(ev["dt"].shape[0]):
        dt = ev["dt"][o].astype(np.float64)
        sel = np.flatnonzero(dt > 1e-9)
        if sel.size == 0:
            continue
        ph = ev["rf_phase"][o, sel].astype(np.float64)
        l1 = _gen_l1(prm, nc, dt[sel], ev["rf_amp"][o, sel].astype(np.float64), np.cos(ph), np.sin(ph),
                     ev["rf_freq"][o, sel].astype(np.float64))
        s = np.where(l1 > 0, np.maximum(0, np.floor(np.log2(np.maximum(l1, 1e-300) / th))), 0).astype(np.int64)
        s = np.minimum(s, 63)
        f = np.where(l1 < lo, cost[3][s], np.where(l1 < hi, cost[5][s], cost[7][s]))
        total += f.sum(axis=1)
    return total


def reference_rule(l1):
    """Reference float32 Pade degree and squarings from the unscaled 1-norm (expm.py:73-77)."""
    s = 0 if l1 <= 0 else max(0, math.floor(math.log2(l1 / 3.925724783138660)))
    m = 3 if l1 < 0.4258730016922831 else 5 if l1 < 1.880152677804762 else 7
    return m, s


def adjoint_flops(prm, nc, events, slots, n_q=2):
    """Useful float64 flops of one bmc_simulate_adjoint call, per voxel of prm [V, S] (summed over rows).

    Per live event: the forward step twice (checkpoint sweep and segment recompute; all P tangents in
    one pass; an event without RF squares its two diagonal blocks apart) and, for each of the n_q RF directions, the second-order columns (Taylor terms of the
    state, P tangents, the RF direction and P mixed terms; squarings by the product rule) and the
    adjoint dot products. Idle lanes of a team are not counted.
    """
    ev = {k: np.asarray(getattr(events, k) if not isinstance(events, dict) else events[k])
          for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    N = 3 * (1 + nc) + 1
    NA = N + 1
    NXY = 2 * (1 + nc)
    FX = 29 * nc + 27
    P = len(slots)
    fdx = sum(dx_flops(sl, nc) for sl in slots)
    dxq = (2 * (5 + 4 * nc), 2 * (3 + 2 * nc))[:n_q]           # rf_amp (rel_b1 pattern), rf_freq (b0 shift)
    term = FX * (1 + P) + fdx + 2 * N * (1 + P)
    mul = NA * 2 * N * N * (1 + 2 * P)
    app = NA * (N + 3 * P * N) + N * (1 + P) * (NA - 1)
    sq2 = NA * 2 * (3 + 6 * P) * N * N
    dots = NA * 2 * (2 + 4 * P) * N
    V = prm.shape[0]
    total = np.zeros(V)
    for o in range(ev["dt"].shape[0]):
        dt = ev["dt"][o].astype(np.float64)
        sel = np.flatnonzero(dt > 1e-9)
        if sel.size == 0:
            continue
        ph = ev["rf_phase"][o, sel].astype(np.float64)
        l1 = _gen_l1(prm, nc, dt[sel], ev["rf_amp"][o, sel].astype(np.float64), np.cos(ph), np.sin(ph),
                     ev["rf_freq"][o, sel].astype(np.float64))
        m, s = degree_and_squarings(l1)
        mulx = np.where(ev["rf_amp"][o, sel] == 0, NA * 2 * (NXY * NXY + (N - NXY) ** 2) * (1 + 2 * P), mul)
        fwd = np.where(s == 0, m * term, NA * m * term + s * mulx + app)
        adj = 0.0
        for dq in dxq:
            term2 = FX * (2 + 2 * P) + fdx + dq + P * dq + fdx + 2 * N * (2 + 2 * P)
            adj = adj + NA * m * term2 + s * sq2 + dots
        total += (2 * fwd + adj).sum(axis=1)
    return total


def single_event_rows(events):
    """The propagating event of every row when the table qualifies for the single-event Jacobian kernel
    (csrc/bmc_cwj.hip; the rule of bmc_plan_create_ex): each row propagates at most one event (dt > 1e-9,
    a delay or RF block), no spoiler or gradient anywhere, and a readout after that event is the row's
    last event. Returns an int array [O] (-1: no propagating event) or None."""
    ev = {k: np.asarray(getattr(events, k) if not isinstance(events, dict) else events[k]) for k in ("type", "dt")}
    typ, dt = ev["type"], ev["dt"].astype(np.float64)
    O, T = typ.shape
    if T == 0 or np.any((typ == 2) | (typ == 4)):
        return None
    live = np.full(O, -1)
    for o in range(O):
        idx = np.flatnonzero(dt[o] > 1e-9)
        if idx.size > 1 or (idx.size == 1 and typ[o, idx[0]] not in (0, 1)):
            return None
        if idx.size == 1:
            live[o] = idx[0]
            adc_after = np.flatnonzero(typ[o, idx[0]:] == 3) + idx[0]
            if np.any(adc_after != T - 1):
                return None
    return live


CWJ_MMAX = 18   # the single-event Jacobian kernel's Taylor degree cap (csrc/bmc_cwj.h)


def reverse_flops(prm, nc, events, slots, live, chunk=None):
    """Useful float64 flops per voxel of prm [V, S] of one single-event Jacobian launch (csrc/bmc_cwj.hip),
    summed over rows, with the kernel's Taylor degree / squaring rule (degree cap CWJ_MMAX):
      the Taylor columns of T            NA m (FX + 2N)
      per squaring level (per lane)      E^2 and E^T G: 2 N^2 each; the rows of E^2 and G E^T: 2 N NA each
      the state                          2 N NA
      per chunk of T <= `chunk` tangents NA m (FX (1 + T) + dX + 2N (1 + T)), then 2 N NA + NA per tangent
    live: single_event_rows(events). chunk defaults to the kernel's: 2, or 1 at four CEST pools."""
    if chunk is None:
        chunk = 1 if nc >= 4 else 2
    ev = {k: np.asarray(getattr(events, k) if not isinstance(events, dict) else events[k])
          for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    N = 3 * (1 + nc) + 1
    NA = N + 1
    FX = 29 * nc + 27
    chunks = [slots[i:i + chunk] for i in range(0, len(slots), chunk)]
    V = prm.shape[0]
    total = np.zeros(V)
    for o in np.flatnonzero(np.asarray(live) >= 0):
        e = [int(live[o])]
        ph = ev["rf_phase"][o, e].astype(np.float64)
        l1 = _gen_l1(prm, nc, ev["dt"][o, e].astype(np.float64), ev["rf_amp"][o, e].astype(np.float64), np.cos(ph),
                     np.sin(ph), ev["rf_freq"][o, e].astype(np.float64))
        m, s = degree_and_squarings(l1, CWJ_MMAX)
        m, s = m[:, 0], s[:, 0]
        f = NA * m * (FX + 2 * N) + s * NA * (4 * N * N + 4 * N * NA) + 2 * N * NA
        for ch in chunks:
            T = len(ch)
            fdx = sum(dx_flops(sl, nc) for sl in ch)
            f = f + NA * m * (FX * (1 + T) + fdx + 2 * N * (1 + T)) + T * (2 * N * NA + NA)
        total += f
    return total
