"""Pulses reused across a row with a shifted RF phase (plan compression level 2), the event-cache
format, and the block structure of events without RF that the kernels' delay products rely on. CPU only:
the rotation identity on the oracle, the host compilation (bmc_plan_compile), the cache reader and
writer, the generator's blocks.

A converted scanner sequence gives every pulse of a train the phase accumulated by the pulses before it
(the reference's sequences/pulseq.py:56-66, 118-130), so no two pulses are equal and a plan can only
reuse one by rotating it: exp(X(phi + d) dt) = R(d) exp(X(phi) dt) R(d)^T, R(d) the rotation by d of
every pool's (x, y). That identity is checked on the oracle's generator (bmc_matrix restates
matrix.py:83-295) and exponential, independently of the kernels.
"""
import os

import numpy as np
import pytest

import bmc_oracle as O
from helpers import four_pool_params

REF_EVENTS = "/root/reference/examples/data/events"
OP_SEQ, OP_BLOCK, OP_KEEP, OP_REUSE = 0, 1, 2, 3


def rot_full(N, nt, d):
    """Rotation by d of every pool's (x, y) on the augmented state (x block [0, nt), y block [nt, 2 nt))."""
    R = np.eye(N + 1)
    c, s = np.cos(d), np.sin(d)
    for i in range(nt):
        R[i, i], R[i, nt + i], R[nt + i, i], R[nt + i, nt + i] = c, -s, s, c
    return R


@pytest.mark.parametrize("d", [0.3, 2.0587284868, -5.1])
def test_a_phase_offset_is_a_rotation_about_z(d):
    p = four_pool_params(k_am=200.0, f_am=3e-3, b0i=0.1, rel_b1=0.9)
    nt = 1 + len(p["cest_k"])
    N = 3 * nt + 1
    for amp, ph, freq, dt in ((101.3, 0.4, -300.0, 5e-4), (12.0, -1.1, 616.0, 0.02), (40.0, 0.0, 0.0, 2.0)):
        A0, C0 = O.bmc_matrix(p, amp, ph, freq)
        A1, C1 = O.bmc_matrix(p, amp, ph + d, freq)
        E0 = O.expm64(O.augmented(A0, C0, dt))
        E1 = O.expm64(O.augmented(A1, C1, dt))
        R = rot_full(N, nt, d)
        np.testing.assert_allclose(R @ O.augmented(A0, C0, dt) @ R.T, O.augmented(A1, C1, dt), rtol=0, atol=1e-12)
        assert np.max(np.abs(R @ E0 @ R.T - E1)) < 1e-13 * max(1.0, np.abs(E1).max())
        # a plausible mistake, the inverse rotation, is not the same thing
        assert np.max(np.abs(R.T @ E0 @ R - E1)) > 1e-6


def _compile(ev, level):
    from bmc_jax import _native
    return _native.compile_program(ev.type, ev.dt, ev.rf_amp, ev.rf_phase, ev.rf_freq, level)


def _pulse_starts(typ_row):
    rf = typ_row == 1
    return np.flatnonzero(rf & ~np.r_[False, rf[:-1]])


def test_scanner_train_keeps_one_pulse_per_row():
    from bmc_jax.sequences import scanner_train
    ev = scanner_train(np.linspace(-5, 5, 7), 1.3, n_pulses=6, n_samples=30, m0_ppm=-300.0)
    ops, rot, rp = _compile(ev, 2)
    for o in range(ev.type.shape[0]):
        row = ops[rp[o]:rp[o + 1]]
        kinds = row[:, 0]
        assert np.sum(kinds == OP_KEEP) == 1 and np.sum(kinds == OP_REUSE) == 5
        starts = _pulse_starts(ev.type[o])
        assert list(row[kinds >= OP_KEEP, 1]) == list(starts)
        assert np.all(row[kinds >= OP_KEEP, 2] == 30)
        ph = ev.rf_phase[o, starts]
        d = ph[1:] - ph[0]
        np.testing.assert_allclose(rot[rp[o]:rp[o + 1]][kinds == OP_REUSE], np.stack([np.cos(d), np.sin(d)], 1),
                                   atol=1e-15)
        # the ops tile the row exactly, in order
        assert row[0, 1] == 0 and np.all(row[1:, 1] == row[:-1, 1] + row[:-1, 2] * np.where(row[:-1, 0] == 1,
                                                                                              row[:-1, 3], 1))
        assert row[-1, 1] + row[-1, 2] == ev.type.shape[1]
    ops1, _, _ = _compile(ev, 1)
    assert np.all(ops1[:, 0] == OP_SEQ)   # no pulse repeats back to back: nothing to chain at level 1
    ops0, _, rp0 = _compile(ev, 0)
    assert len(ops0) == ev.type.shape[0] and np.all(ops0[:, 2] == ev.type.shape[1])


def test_reuse_needs_one_phase_offset_for_the_whole_pulse():
    from bmc_jax.sequences import scanner_train
    ev = scanner_train(np.array([-2.0, 3.0]), 1.0, n_pulses=5, n_samples=20)
    ph = np.array(ev.rf_phase, copy=True)
    s = _pulse_starts(ev.type[0])
    ph[0, s[2] + 7] += 1e-6                       # one sample of pulse 2 in row 0 off the common offset
    ev = ev._replace(rf_phase=ph)
    ops, rot, rp = _compile(ev, 2)
    r0, r1 = ops[rp[0]:rp[1]], ops[rp[1]:rp[2]]
    assert s[2] not in r0[r0[:, 0] >= OP_KEEP, 1]          # pulse 2 runs event by event
    assert np.sum(r0[:, 0] == OP_REUSE) == 3 and np.sum(r1[:, 0] == OP_REUSE) == 4
    amp = np.array(ev.rf_amp, copy=True)
    amp[1, s[3] + 1] *= 1.0 + 1e-12                          # any amplitude change breaks the match too
    ops, _, rp = _compile(ev._replace(rf_amp=amp), 2)
    assert np.sum(ops[rp[1]:rp[2], 0] == OP_REUSE) == 3


def test_identical_back_to_back_pulses_still_chain_as_blocks():
    """The level-1 runs take precedence: a train of identical pulse + gap units is one BLOCK."""
    from bmc_jax.sequences import pulsed_train
    ev = pulsed_train(np.array([1.0, -2.0]), n_pulses=8, n_samples=12)
    for level in (1, 2):
        ops, _, rp = _compile(ev, level)
        kinds = ops[:, 0]
        assert np.sum(kinds == OP_BLOCK) == 2 and np.all(ops[kinds == OP_BLOCK, 3] == 8)
        assert not np.any(kinds >= OP_KEEP)


def test_event_cache_round_trip(tmp_path):
    from bmc_jax.sequences import load_events, save_events, scanner_train
    ppm = np.r_[-300.0, np.linspace(-5, 5, 11)]
    ev = scanner_train(ppm[1:], 0.8, n_pulses=3, n_samples=10, m0_ppm=-300.0)
    f = tmp_path / "rung_events.npz"
    save_events(f, ev, ppm, max_pulse_samples=10, b1_uT=0.8)
    back, meta = load_events(f, dtype=np.float64, return_meta=True)
    assert back.type.shape == (11, ev.type.shape[1])
    for k in ev._fields:
        np.testing.assert_array_equal(np.asarray(getattr(back, k)), np.asarray(getattr(ev, k))[1:])
    np.testing.assert_array_equal(meta["offsets_ppm"], ppm[1:])
    assert int(meta["max_pulse_samples"]) == 10 and float(meta["b1_uT"]) == 0.8
    keep = np.abs(ppm[1:]) >= 1.0
    sub, meta = load_events(f, keep=keep, return_meta=True)
    assert sub.type.shape[0] == keep.sum() and sub.dt.dtype == np.float32
    np.testing.assert_array_equal(meta["offsets_ppm"], ppm[1:][keep])
    np.savez(tmp_path / "bad.npz", dt=np.zeros((2, 3)))
    with pytest.raises(ValueError, match="not an event cache"):
        load_events(tmp_path / "bad.npz")


@pytest.mark.skipif(not os.path.isdir(REF_EVENTS), reason="the reference's event caches are not mounted")
def test_the_reference_caches_compile_to_kept_pulses():
    """The shipped scanner conversions (examples/data/events/*.npz): every pulse after a row's first is
    applied as a rotated copy of it."""
    from bmc_jax.sequences import load_events
    for name in sorted(os.listdir(REF_EVENTS)):
        ev, meta = load_events(os.path.join(REF_EVENTS, name), return_meta=True)
        assert int(meta["max_pulse_samples"]) == 200
        assert ev.type.shape[0] == len(meta["offsets_ppm"]) == 101
        ops, rot, rp = _compile(ev, 2)
        for o in (0, 50, 100):
            row = ops[rp[o]:rp[o + 1]]
            n_pulses = len(_pulse_starts(np.asarray(ev.type[o])))
            assert np.sum(row[:, 0] == OP_KEEP) == 1, name
            assert np.sum(row[:, 0] == OP_REUSE) == n_pulses - 1, name


def _sym_train(ppm, phase=0.0):
    """Three Gaussian pulses of 64 symmetric samples with gaps, RF at `phase` (oracle's pulsed_train)."""
    ev = dict(O.pulsed_train(np.asarray(ppm, float), b1_peak_uT=2.0, n_pulses=3, pulse_s=0.02, n_samples=64,
                             gap_s=0.02, recovery_s=1.0, b0_T=3.0, dtype=np.float64))
    ev["rf_phase"] = np.where(ev["type"] == 1, phase, 0.0)
    return ev


def test_plan_marks_only_palindromes():
    """Host compilation: the half of a symmetric pulse at one phase, none for a phase that changes inside
    the pulse or a pulse that is not symmetric; level 2 marks nothing."""
    from bmc_jax import _native
    ev = _sym_train([1.0, 2.0])
    ops, rot, rp = _native.compile_program(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"], 3)
    blk = ops[ops[:, 0] == 1]
    assert np.all(blk[:, 4] == 32) and np.allclose(rot[ops[:, 0] == 1], [1.0, 0.0])
    ev7 = _sym_train([1.0], phase=0.7)
    ops, rot, _ = _native.compile_program(ev7["type"], ev7["dt"], ev7["rf_amp"], ev7["rf_phase"], ev7["rf_freq"], 3)
    assert np.allclose(rot[ops[:, 4] > 0], [np.cos(0.7), np.sin(0.7)])
    ops2, _, _ = _native.compile_program(ev["type"], ev["dt"], ev["rf_amp"], ev["rf_phase"], ev["rf_freq"], 2)
    assert np.all(ops2[:, 4] == 0)
    chirp = dict(ev7)
    chirp["rf_phase"] = np.where(ev7["type"] == 1, np.cumsum(ev7["type"] == 1, axis=1) * 1e-3, 0.0)
    ops, _, _ = _native.compile_program(chirp["type"], chirp["dt"], chirp["rf_amp"], chirp["rf_phase"],
                                        chirp["rf_freq"], 3)
    assert np.all(ops[:, 4] == 0)
    skew = dict(ev)
    amp = skew["rf_amp"].copy()
    amp[:, [2, 67, 132]] *= 1.0 + 1e-12      # the second sample of every pulse: still repeated, not symmetric
    skew["rf_amp"] = amp
    ops, _, _ = _native.compile_program(skew["type"], skew["dt"], skew["rf_amp"], skew["rf_phase"],
                                        skew["rf_freq"], 3)
    assert np.any(ops[:, 0] == 1) and np.all(ops[:, 4] == 0)


def sym_matrix(p, phi, flip_y=True):
    """S of the palindrome factorisation at RF phase phi: R(phi) S0 R(-phi), S0 = diag over (x, y, z) of
    water and each pool of (1, f_i) with the y block negated, and f_mt on MT's z (flip_y=False is a
    plausible mistake that drops the sign)."""
    f = np.asarray(p["cest_f"], float)
    nt = 1 + f.size
    N = 3 * nt + 1
    sc = np.concatenate([[1.0], f])
    s0 = np.concatenate([sc, (-1.0 if flip_y else 1.0) * sc, sc, [p["mt_f"]]])
    R = rot_full(N, nt, phi)[:N, :N]
    return R @ np.diag(s0) @ R.T


@pytest.mark.parametrize("phi", [0.0, 0.7, -2.3])
def test_the_generator_is_symmetrisable_and_a_palindrome_is_its_transpose(phi):
    """The identity behind plan level 3 (mirror_chain in csrc/bmc_device.h), on the oracle's generator and
    exponential: S^-1 A S = A^T for every event, hence a chained run E_{2h-1} .. E_0 that reads the same
    backwards equals Q^R Q with Q^R's linear part S Q_A^T S^-1 and its affine column S w, w the affine
    column of exp(X'_0) .. exp(X'_{h-1}) for X' = [[A^T, S^-1 C], [0, 0]]. Without the y sign the
    symmetrisation fails."""
    p = four_pool_params(k_am=300.0, f_am=2e-3, b0i=0.05, rel_b1=1.1)
    N = 3 * (1 + len(p["cest_k"])) + 1
    S = sym_matrix(p, phi)
    Si = np.linalg.inv(S)
    amps = [3.0, 19.0, 47.5, 80.2]                 # the first half of a symmetric (Gaussian-like) pulse
    for amp in amps:
        A, _ = O.bmc_matrix(p, amp, phi, 420.0)
        assert np.max(np.abs(Si @ A @ S - A.T)) < 1e-12 * np.max(np.abs(A))
        Sb = sym_matrix(p, phi, flip_y=False)
        assert np.max(np.abs(np.linalg.inv(Sb) @ A @ Sb - A.T)) > 1e-3 * np.max(np.abs(A))
    dt = 1e-4
    Es = [O.expm64(O.augmented(*O.bmc_matrix(p, a, phi, 420.0), dt)) for a in amps]
    h = len(Es)
    direct = np.eye(N + 1)
    for E in Es + Es[::-1]:                        # E_0 .. E_{h-1}, then E_{h-1} .. E_0
        direct = E @ direct
    Q = np.eye(N + 1)
    for E in Es:
        Q = E @ Q
    Wp = np.eye(N + 1)                             # exp(X'_0) .. exp(X'_{h-1}) from the transposed generators
    for a in amps:
        A, C = O.bmc_matrix(p, a, phi, 420.0)
        Xp = np.zeros((N + 1, N + 1))
        Xp[:N, :N], Xp[:N, N] = A.T, Si @ C
        Wp = Wp @ O.expm64(Xp * dt)
    QR = np.eye(N + 1)
    QR[:N, :N] = S @ Q[:N, :N].T @ Si
    QR[:N, N] = S @ Wp[:N, N]
    assert np.max(np.abs(QR @ Q - direct)) < 1e-13
    assert h == 4


FIELDS = [("water_r1", None), ("water_r2", None), ("water_f", None), ("mt_k", None), ("mt_f", None),
          ("mt_dw", None), ("mt_r1", None), ("mt_r2", None), ("b0", None), ("gamma", None), ("b0_inhom", None),
          ("rel_b1", None), ("cest_k", 0), ("cest_f", 0), ("cest_dw", 0), ("cest_r1", 0), ("cest_r2", 0),
          ("cest_k", 1), ("cest_f", 1), ("cest_dw", 1), ("cest_r1", 1), ("cest_r2", 1)]


def _off_blocks(X, nt):
    """The two off-diagonal blocks of an augmented matrix split into the transverse rows [0, 2 nt) and
    the z rows with MT's z and the homogeneous row [2 nt, N + 1)."""
    xy = slice(0, 2 * nt)
    z = slice(2 * nt, X.shape[0])
    return np.concatenate([X[xy, z].ravel(), X[z, xy].ravel()])


@pytest.mark.parametrize("lineshape", ["superlorentzian", "lorentzian"])
def test_an_event_without_rf_is_block_diagonal(lineshape):
    """The kernels square a delay's propagator and multiply it into a chained one block by block
    (rf_free / mul_cols' BD form in csrc/bmc_device.h). That rests on this: without RF the generator
    (matrix.py:83-295, restated as bmc_matrix) and its derivative with respect to every packed field have
    no entry between the transverse rows and the z rows, so the exponential and its Frechet derivatives
    have none either. With RF both the generator and the B1 derivative couple the blocks (the plausible
    mistake of taking an RF event for a delay)."""
    p = four_pool_params(k_am=1200.0, f_am=5e-3, b0i=-0.2, rel_b1=0.9, lineshape=lineshape)
    nt = 1 + len(p["cest_k"])
    for freq, dt in ((-300.0, 0.02), (616.0, 3.0), (0.0, 1e-3)):
        X = O.augmented(*O.bmc_matrix(p, 0.0, 0.0, freq), dt)
        assert np.all(_off_blocks(X, nt) == 0.0)
        assert np.abs(X).sum() > 0
        E = O.expm64(X)
        assert np.max(np.abs(_off_blocks(E, nt))) < 1e-15
        for field, idx in FIELDS:
            v = p[field] if idx is None else p[field][idx]
            h = 1e-6 * max(abs(v), 1e-3)
            Xp = O.augmented(*O.bmc_matrix(O.write_params(p, [v + h], [(field, idx)]), 0.0, 0.0, freq), dt)
            Xm = O.augmented(*O.bmc_matrix(O.write_params(p, [v - h], [(field, idx)]), 0.0, 0.0, freq), dt)
            assert np.all(_off_blocks(Xp, nt) == 0.0) and np.all(_off_blocks(Xm, nt) == 0.0), field
        # with RF the blocks are coupled, and so is the rel_b1 derivative
        Xr = O.augmented(*O.bmc_matrix(p, 40.0, 0.3, freq), dt)
        assert np.max(np.abs(_off_blocks(Xr, nt))) > 1e-3
        Xr2 = O.augmented(*O.bmc_matrix(O.write_params(p, [1.0], [("rel_b1", None)]), 40.0, 0.3, freq), dt)
        assert np.max(np.abs(_off_blocks(Xr2 - Xr, nt))) > 1e-4
