"""CPU oracle for the simulate-and-fit path: a float64 NumPy restatement of BMC-JAX.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may use it, and only as the
checker (or the timed CPU baseline), never as the thing measured or shipped.

Every function cites the reference file:line it restates (paths relative to the BMC-JAX tree,
`src/bmc_jax/...`). The reference itself cannot be imported here (JAX is absent), so this oracle
is pinned by closed forms, invariants and scipy (see tests/test_oracle.py):

  bmc_matrix         pinned by: pure-relaxation closed form, steady state -A^-1 C, the far-wing
                     value water_f, an independent SuperLorentzian quadrature, Lorentzian formula.
  expm / propagate   pinned by: scipy.linalg.expm (Al-Mohy--Higham) vs the oracle's own Pade-13
                     (Higham 2005) in C, closed-form 2x2 rotation-relaxation blocks, dt=0 identity.
  simulate_sequence  pinned by: the reference's exact-equality tests (front padding, stacking,
                     tail-padding trap) and the physics invariants of tests/test_solver.py.
  jacobian           pinned by: two independent routes (central differences of Z; scipy's
                     expm_frechet chained with a differenced generator) that must agree.
  levenberg_marquardt pinned by: Rosenbrock -> (1, 1) and the held-entry rule (test_fitting.py,
                     test_joint.py of the reference).
  lm_trial (active bounds) pinned by: the reduced-problem solve (test_joint.py:125-143).

Parity unpinned: bit-level float32 outputs of jax.scipy.linalg.expm and JAX's own float64 Pade-13
rounding cannot be checked without JAX; the oracle's float64 agrees with scipy to ~1e-14.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import scipy.linalg

# sequences/saturation.py:18-20
GAMMA_RAD_PER_S_UT = 267.5153
GAMMA_HZ_PER_T = 42.5764e6

DELAY, RF_PULSE, GRADIENT, ADC, SPOILER = 0, 1, 2, 3, 4          # simulation/events.py:12-18

FIELDS = ("water_r1", "water_r2", "water_f", "cest_k", "cest_f", "cest_dw", "cest_r1", "cest_r2",
          "mt_k", "mt_f", "mt_dw", "mt_r1", "mt_r2", "mt_lineshape", "b0", "gamma", "b0_inhom",
          "rel_b1")                                                   # simulation/params.py:18-49


def make_params(water_t1, water_t2, pools=(), mt=None, b0=3.0, gamma=GAMMA_RAD_PER_S_UT,
                b0_inhom=0.0, rel_b1=1.0, water_f=1.0, n_cest=None):
    """A parameter record (dict of float64 values) as `to_sim_params` builds it.

    pools: sequence of dicts with f, k, dw, t1, t2. mt: dict with f, k, dw, t1, t2, lineshape.
    Follows simulation/params.py:51-152 (zero padding :86-94, the no-MT defaults :117-125).
    """
    nc = len(pools) if n_cest is None else n_cest
    pad = nc - len(pools)
    if pad < 0:
        raise ValueError("n_cest smaller than the number of pools")
    p = dict(water_r1=1.0 / water_t1, water_r2=1.0 / water_t2, water_f=water_f,
             cest_k=np.array([q["k"] for q in pools] + [0.0] * pad, float),
             cest_f=np.array([q["f"] for q in pools] + [0.0] * pad, float),
             cest_dw=np.array([q["dw"] for q in pools] + [0.0] * pad, float),
             cest_r1=np.array([1.0 / q["t1"] for q in pools] + [0.0] * pad, float),
             cest_r2=np.array([1.0 / q["t2"] for q in pools] + [0.0] * pad, float),
             b0=b0, gamma=gamma, b0_inhom=b0_inhom, rel_b1=rel_b1)
    if mt is None:
        p.update(mt_k=0.0, mt_f=0.0, mt_dw=0.0, mt_r1=0.0, mt_r2=1.0, mt_lineshape=0)
    else:
        ls = {"lorentzian": 1, "superlorentzian": 2}.get(str(mt.get("lineshape", "superlorentzian")).lower(), 0)
        p.update(mt_k=mt["k"], mt_f=mt["f"], mt_dw=mt["dw"], mt_r1=1.0 / mt["t1"], mt_r2=1.0 / mt["t2"],
                 mt_lineshape=ls)
    return p


def n_state(p):
    """State length 3 * (1 + n_cest) + 1 (simulation/params.py:189-191; matrix.py:216-219)."""
    return 3 * (1 + len(p["cest_k"])) + 1


def mz_loc(p):
    """Index of water Mz: 2 * (1 + n_cest) (io/parameters/parameters.py:57-67)."""
    return 2 * (1 + len(p["cest_k"]))


def initial_state(p, m_start=1.0):
    """simulation/params.py:154-212: water_f, cest_f, mt_f on the z slots, all times m_start."""
    nc = len(p["cest_k"])
    nt = 1 + nc
    M = np.zeros(3 * nt + 1)
    M[2 * nt] = p["water_f"]
    M[2 * nt + 1:3 * nt] = p["cest_f"]
    M[3 * nt] = p["mt_f"]
    return M * m_start


# ---- lineshapes: simulation/matrix.py:24-81 -------------------------------------------------------
_SL_U = np.linspace(0.0, 1.0, 101)
_SL_DU = 0.01
_SL_T1 = np.maximum(np.abs(3 * _SL_U ** 2 - 1), 1e-6)


def interpolate_sl(delta, t2):
    """matrix.py:34-38: 101-point rectangle rule over u = cos(theta), times pi * du."""
    arg = delta * t2 / _SL_T1
    integrand = np.sqrt(2 / np.pi) * t2 / _SL_T1 * np.exp(-2 * arg ** 2)
    return float(np.sum(integrand) * np.pi * _SL_DU)


def interpolate_chs(delta, t2, w0):
    """matrix.py:40-60: cubic Hermite bridge over |delta| < w0, with the '+1' offsets of :53."""
    px = [-300 - w0, -100 - w0, 100 + w0, 300 + w0]
    py = [interpolate_sl(x, t2) for x in px]
    d0y = 30.0 * (py[1] - py[0])
    d1y = 30.0 * (py[3] - py[2])
    c = abs((delta - px[1] + 1) / (px[2] - px[1] + 1))
    h0 = 2 * c ** 3 - 3 * c ** 2 + 1
    h1 = -2 * c ** 3 + 3 * c ** 2
    h2 = c ** 3 - 2 * c ** 2 + c
    h3 = c ** 3 - c ** 2
    return h0 * py[1] + h1 * py[2] + h2 * d0y + h3 * d1y


def super_lorentzian(delta, t2, w0):
    """matrix.py:62-68."""
    return interpolate_sl(delta, t2) if abs(delta) >= w0 else interpolate_chs(delta, t2, w0)


def lorentzian(delta, t2):
    """matrix.py:70-81 (no 1/pi)."""
    return t2 / (1 + (delta * t2) ** 2)


# ---- the system matrix: simulation/matrix.py:83-295 -----------------------------------------------
def bmc_matrix(p, rf_amp, rf_phase, rf_freq):
    """A (N x N) and C (N,) of dM/dt = A M + C for one event, exactly as matrix.py:83-295."""
    r1w, r2w = p["water_r1"], p["water_r2"]
    k, f = np.asarray(p["cest_k"], float), np.asarray(p["cest_f"], float)
    r1c, r2c, dwc = np.asarray(p["cest_r1"], float), np.asarray(p["cest_r2"], float), np.asarray(p["cest_dw"], float)
    kmt, fmt, r1mt, r2mt, dwmt = p["mt_k"], p["mt_f"], p["mt_r1"], p["mt_r2"], p["mt_dw"]
    kwc = k * f
    kwmt = kmt * fmt
    w1 = 2 * np.pi * rf_amp * p["rel_b1"]                                  # :121-123
    w1x, w1y = w1 * np.cos(rf_phase), w1 * np.sin(rf_phase)
    larmor = p["gamma"] * 1e6 * p["b0"]                                    # :127-128
    dw0 = larmor * p["b0_inhom"] * 1e-6                                    # :131
    eff = 2 * np.pi * rf_freq - dw0                                        # :162
    off_c = dwc * 1e-6 * larmor - eff                                      # :165-166
    off_w = eff                                                            # :169
    off_mt = eff - dwmt * 1e-6 * larmor                                    # :172-173
    t2mt = 1.0 / r2mt
    g = 0.0
    if fmt > 1e-9:                                                         # :203-208
        ls = int(p["mt_lineshape"])
        if ls == 1:
            g = lorentzian(off_mt, t2mt)
        elif ls == 2:
            g = super_lorentzian(off_mt, t2mt, larmor * 1e-6)
    w_mt = w1 ** 2 * g if fmt > 1e-9 else 0.0                              # :210-213

    nc = k.size
    nt = 1 + nc
    N = 3 * nt + 1
    A = np.zeros((N, N))
    xw, yw, zw = 0, nt, 2 * nt
    xc = np.arange(1, nt)
    yc = np.arange(nt + 1, 2 * nt)
    zc = np.arange(2 * nt + 1, 3 * nt)
    zmt = 3 * nt
    ks = kwc.sum()
    A[xw, xw] = -r2w - ks; A[xw, yw] = off_w; A[xw, zw] = -w1y              # :238-241
    A[yw, yw] = -r2w - ks; A[yw, xw] = -off_w; A[yw, zw] = w1x              # :244-246
    A[zw, zw] = -r1w - ks - kwmt; A[zw, yw] = -w1x; A[zw, xw] = w1y         # :249-251
    A[xc, xc] = -r2c - k; A[xc, yc] = -off_c; A[xc, zc] = -w1y              # :255-257
    A[yc, yc] = -r2c - k; A[yc, xc] = off_c; A[yc, zc] = w1x                # :260-262
    A[zc, zc] = -r1c - k; A[zc, xc] = w1y; A[zc, yc] = -w1x                 # :265-267
    A[xw, xc] = k; A[yw, yc] = k; A[zw, zc] = k                             # :271-273
    A[xc, xw] = kwc; A[yc, yw] = kwc; A[zc, zw] = kwc                       # :276-278
    A[zmt, zmt] = -r1mt - kmt - w_mt; A[zmt, zw] = kwmt; A[zw, zmt] = kmt   # :282-286
    C = np.zeros(N)
    C[zw] = r1w * 1.0; C[zc] = r1c * f; C[zmt] = r1mt * fmt                 # :289-293
    return A, C


def augmented(A, C, dt):
    """simulation/expm.py:98-106: [[A, C], [0, 0]] * dt."""
    N = A.shape[0]
    X = np.zeros((N + 1, N + 1))
    X[:N, :N] = A
    X[:N, N] = C
    return X * dt


# ---- the reference's float32 expm, restated (simulation/expm.py:50-83) -----------------------------
def ref_squarings(X, dtype=np.float32):
    """Squaring count of the reference rule, expm.py:73-74 (float32) -- also the NaN rule (> 16)."""
    l1 = float(np.max(np.sum(np.abs(np.asarray(X, dtype)), axis=0)))
    if l1 == 0.0:
        return 0
    return int(max(0.0, math.floor(math.log2(l1 / 3.925724783138660))))


def expm_ref32(X, max_squarings=16):
    """expm.py:50-83 in float32: Pade 3/5/7 by the unscaled norm, Gauss-Jordan, squarings.

    Used only to measure what the reference's own float32 algorithm errs by; not a parity target.
    """
    A = np.asarray(X, np.float32)
    n = A.shape[0]
    I = np.eye(n, dtype=np.float32)
    l1 = np.float32(np.max(np.sum(np.abs(A), axis=0)))
    n_sq = max(0.0, math.floor(math.log2(l1 / np.float32(3.925724783138660)))) if l1 > 0 else 0.0
    As = A / np.float32(2.0 ** n_sq)
    A2 = As @ As
    if l1 < 4.258730016922831e-001:
        b = (120., 60., 12., 1.)
        U = As @ (b[3] * A2 + b[1] * I); V = b[2] * A2 + b[0] * I
    elif l1 < 1.880152677804762e+000:
        b = (30240., 15120., 3360., 420., 30., 1.)
        A4 = A2 @ A2
        U = As @ (b[5] * A4 + b[3] * A2 + b[1] * I); V = b[4] * A4 + b[2] * A2 + b[0] * I
    else:
        b = (17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.)
        A4 = A2 @ A2; A6 = A4 @ A2
        U = As @ (b[7] * A6 + b[5] * A4 + b[3] * A2 + b[1] * I)
        V = b[6] * A6 + b[4] * A4 + b[2] * A2 + b[0] * I
    U = U.astype(np.float32); V = V.astype(np.float32)
    R = np.linalg.solve(V - U, U + V).astype(np.float32)
    for _ in range(int(n_sq)):
        R = (R @ R).astype(np.float32)
    if n_sq > max_squarings:
        R = np.full_like(R, np.nan)
    return R


def expm64(X):
    """Float64 matrix exponential (scipy: Al-Mohy--Higham scaling and squaring)."""
    return scipy.linalg.expm(X)


# ---- the scan: simulation/simulate.py:22-148 -------------------------------------------------------
def simulate_sequence(p, ev, M_init, reset_after_adc=True, expm=expm64, return_trajectory=False,
                      nan_rule=True):
    """One offset row: simulate.py:80-132 with simulate_step :22-75.

    ev: dict of 1-D arrays type, dt, rf_amp, rf_phase, rf_freq. Returns last_M (or the trajectory).
    nan_rule: an event whose reference squaring count exceeds 16 yields NaN (expm.py:83).
    """
    nt = 1 + len(p["cest_k"])
    M = np.array(M_init, float)
    last = M.copy()
    traj = []
    for i in range(len(ev["dt"])):
        typ = int(ev["type"][i]); dt = float(ev["dt"][i])
        if dt > 1e-9:                                                      # simulate.py:51
            A, C = bmc_matrix(p, float(ev["rf_amp"][i]), float(ev["rf_phase"][i]), float(ev["rf_freq"][i]))
            X = augmented(A, C, dt)
            if nan_rule and ref_squarings(X) > 16:
                M = np.full_like(M, np.nan)
            else:
                E = expm(X)
                M = (E @ np.append(M, 1.0))[:-1]
        if typ in (GRADIENT, SPOILER):                                     # simulate.py:55-72
            M = M.copy(); M[:2 * nt] = 0.0
        last = M.copy()
        traj.append(last)
        if typ == ADC and reset_after_adc:                                 # simulate.py:105-114
            M = np.array(M_init, float)
    return np.array(traj) if return_trajectory else last


def simulate_spectrum(p, events, M_init, **kw):
    """simulate.py:134-148: every offset row of events (dict of (O, T) arrays) -> (O, N)."""
    O = np.asarray(events["dt"]).shape[0]
    rows = [{k: np.asarray(v)[o] for k, v in events.items()} for o in range(O)]
    return np.array([simulate_sequence(p, r, M_init, **kw) for r in rows])


# ---- writing free values: fitting/params.py:100-112 ------------------------------------------------
def write_params(p, x_phys, slots):
    """slots: list of (field, index). Returns a new record with x_phys written in."""
    q = {k: (np.array(v, float) if isinstance(v, np.ndarray) else v) for k, v in p.items()}
    for (field, idx), val in zip(slots, x_phys):
        if idx is None:
            q[field] = float(val)
        else:
            q[field] = np.array(q[field], float)
            q[field][idx] = float(val)
    return q


def z_of(p, events, M_init, mz, slots, x_phys):
    return simulate_spectrum(write_params(p, x_phys, slots), events, M_init)[:, mz]


def jacobian_fd(p, events, M_init, mz, slots, x_phys, rel_h=1e-5):
    """Central differences of Z w.r.t. the physical values (float64)."""
    x = np.asarray(x_phys, float)
    cols = []
    for j in range(x.size):
        h = rel_h * max(abs(x[j]), 1e-3)
        xp, xm = x.copy(), x.copy()
        xp[j] += h; xm[j] -= h
        cols.append((z_of(p, events, M_init, mz, slots, xp) - z_of(p, events, M_init, mz, slots, xm)) / (2 * h))
    return np.stack(cols, axis=1)


def jacobian_frechet(p, events, M_init, mz, slots, x_phys, rel_h=1e-6):
    """Forward-mode Jacobian by the chain rule: dE = expm_frechet(X, dX) with dX differenced.

    An independent route from jacobian_fd: the exponential is differentiated exactly (scipy's
    Al-Mohy--Higham Frechet derivative); only the generator's dependence on a value is differenced.
    """
    x = np.asarray(x_phys, float)
    q = write_params(p, x, slots)
    nt = 1 + len(p["cest_k"])
    O = np.asarray(events["dt"]).shape[0]
    P = x.size
    J = np.zeros((O, P))
    for o in range(O):
        ev = {k: np.asarray(v)[o] for k, v in events.items()}
        M = np.array(M_init, float)
        dM = np.zeros((P, M.size))
        lastdM = dM.copy()
        for i in range(len(ev["dt"])):
            typ = int(ev["type"][i]); dt = float(ev["dt"][i])
            if dt > 1e-9:
                args = (float(ev["rf_amp"][i]), float(ev["rf_phase"][i]), float(ev["rf_freq"][i]))
                A, C = bmc_matrix(q, *args)
                X = augmented(A, C, dt)
                E = expm64(X)
                Ma = np.append(M, 1.0)
                newdM = np.zeros_like(dM)
                for j in range(P):
                    h = rel_h * max(abs(x[j]), 1e-3)
                    xp, xm = x.copy(), x.copy(); xp[j] += h; xm[j] -= h
                    Ap, Cp = bmc_matrix(write_params(p, xp, slots), *args)
                    Am, Cm = bmc_matrix(write_params(p, xm, slots), *args)
                    dX = (augmented(Ap, Cp, dt) - augmented(Am, Cm, dt)) / (2 * h)
                    L = scipy.linalg.expm_frechet(X, dX, compute_expm=False)
                    newdM[j] = (L @ Ma)[:-1] + (E @ np.append(dM[j], 0.0))[:-1]
                dM = newdM
                M = (E @ Ma)[:-1]
            if typ in (GRADIENT, SPOILER):
                M = M.copy(); M[:2 * nt] = 0.0; dM = dM.copy(); dM[:, :2 * nt] = 0.0
            lastdM = dM.copy()
            if typ == ADC:
                M = np.array(M_init, float); dM = np.zeros_like(dM)
        J[o] = lastdM[:, mz]
    return J


# ---- sequences -------------------------------------------------------------------------------------
def hz_per_ppm(b0_T):
    """saturation.py:23-25."""
    return GAMMA_HZ_PER_T * b0_T / 1e6


def cw_saturation(ppm, b1_uT, tsat_s, b0_T=2.8936, m0_ppm=None):
    """saturation.py:28-84: (O, 2) = [RF block, ADC]; float32 values as the reference stores them."""
    ppm = np.atleast_1d(np.asarray(ppm, np.float64))
    if m0_ppm is not None:
        ppm = np.concatenate([[float(m0_ppm)], ppm])
    n = ppm.size
    amp = b1_uT * GAMMA_RAD_PER_S_UT / (2 * np.pi)
    freq = ppm * hz_per_ppm(b0_T)
    col = lambda a, b: np.stack([np.broadcast_to(np.asarray(a, np.float64), (n,)),
                                 np.broadcast_to(np.asarray(b, np.float64), (n,))], axis=1)
    return dict(type=col(RF_PULSE, ADC).astype(np.int32), dt=col(tsat_s, 0.0).astype(np.float32),
                rf_amp=col(amp, 0.0).astype(np.float32), rf_phase=col(0.0, 0.0).astype(np.float32),
                rf_freq=col(freq, np.zeros(n)).astype(np.float32),
                adc_idx=col(-1, 0).astype(np.int32))


def pulsed_train(ppm, b1_peak_uT=2.0, n_pulses=50, pulse_s=0.02, n_samples=200, gap_s=0.02,
                 recovery_s=3.0, b0_T=3.0, dtype=np.float32):
    """[DELAY recovery] + n_pulses x ([n_samples Gaussian RF steps] + [DELAY gap]) + [SPOILER] + [ADC].

    Gaussian sigma = pulse/6, sampled at step centres; delays carry rf_freq 0 as the reference's
    Pulseq conversion writes them. 1 + n_pulses (n_samples + 1) + 2 events per offset row.
    """
    ppm = np.atleast_1d(np.asarray(ppm, np.float64))
    O = ppm.size
    peak = b1_peak_uT * GAMMA_RAD_PER_S_UT / (2 * np.pi)
    dts = pulse_s / n_samples
    tc = (np.arange(n_samples) + 0.5) * dts
    sig = pulse_s / 6.0
    shape = peak * np.exp(-0.5 * ((tc - pulse_s / 2) / sig) ** 2)
    typ = [DELAY] + ([RF_PULSE] * n_samples + [DELAY]) * n_pulses + [SPOILER, ADC]
    dt = [recovery_s] + ([dts] * n_samples + [gap_s]) * n_pulses + [0.0, 0.0]
    amp = [0.0] + (list(shape) + [0.0]) * n_pulses + [0.0, 0.0]
    isrf = np.array([t == RF_PULSE for t in typ])
    T = len(typ)
    freq = np.where(isrf[None, :], ppm[:, None] * hz_per_ppm(b0_T), 0.0)
    rep = lambda a: np.broadcast_to(np.asarray(a, np.float64)[None, :], (O, T))
    adc = np.where(np.array(typ) == ADC, 0, -1)
    return dict(type=rep(typ).astype(np.int32), dt=rep(dt).astype(dtype), rf_amp=rep(amp).astype(dtype),
                rf_phase=np.zeros((O, T), dtype), rf_freq=freq.astype(dtype), adc_idx=rep(adc).astype(np.int32))


# ---- Levenberg-Marquardt: fitting/lm.py:20-67 ------------------------------------------------------
def levenberg_marquardt(resid, jac, x0, lo, hi, n_iter=40, lam0=1e-2, free=None, trajectory=False):
    """lm.py:43-61 in float64: g = J^T r, H = J^T J, A = H + lam (diag H + 1e-9 I), solve, clip,
    accept iff strictly better (lam x 0.5), else reject (lam x 3). Fixed n_iter, no early stop."""
    x = np.asarray(x0, float).copy()
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    lam = float(lam0)
    P = x.size
    X, Cs = [], []
    for _ in range(n_iter):
        r = np.asarray(resid(x), float)
        c = float(np.sum(r ** 2))
        J = np.asarray(jac(x), float)
        if free is not None:
            J = J * free
        g = J.T @ r
        H = J.T @ J
        A = H + lam * (np.diag(np.diag(H)) + 1e-9 * np.eye(P))
        if free is not None:
            A = A + np.diag(1.0 - np.asarray(free, float))
        delta = np.linalg.solve(A, -g)
        xn = np.clip(x + delta, lo, hi)
        cn = float(np.sum(np.asarray(resid(xn), float) ** 2))
        better = cn < c
        if better:
            x, lam = xn, lam * 0.5
        else:
            lam = lam * 3.0
        X.append(x.copy()); Cs.append(cn if better else c)
    return (x, np.array(X), np.array(Cs)) if trajectory else x


def lm_trial(x, lam, J, r, lo, hi, active_bounds=False):
    """fitting/joint.py:149-162: the clipped trial point, optionally holding railed values."""
    g, H = J.T @ r, J.T @ J
    P = x.size
    if active_bounds:
        keep = np.where(((x <= lo) & (g > 0)) | ((x >= hi) & (g < 0)), 0.0, 1.0)
        H = H * keep[:, None] * keep[None, :] + np.diag(1.0 - keep)
        g = g * keep
    return np.clip(x + np.linalg.solve(H + lam * (np.diag(np.diag(H)) + 1e-9 * np.eye(P)), -g), lo, hi)


def as_namespace(p):
    return SimpleNamespace(**p)
