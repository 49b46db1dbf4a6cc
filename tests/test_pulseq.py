"""The Pulseq reader and converter (src/bmc_jax/sequences/pulseq.py:20-428 of the reference).

Pinned three ways: the reference's own scanner sequences against the event caches it converted them to
(examples/data/seqs/*.seq -> examples/data/events/*.npz, bit for bit, and the md5 signature each file
carries); shape decoding against hand-expanded cases; the conversion rules against sequences written here
whose events follow by hand (block-pulse steps, zero samples, decimation, phase accumulation).
"""
import math
import os

import numpy as np
import pytest

from bmc_jax.sequences import cw_saturation, hz_per_ppm
from bmc_jax.sequences.pulseq import (SeqWriter, decode_shape, encode_shape, events_from_seq,
                                      events_from_seq_batch, read_seq)
from bmc_jax.simulation.events import EventType as E

PHANTOM = os.path.join(os.path.dirname(__file__), "golden", "phantom")
SEQS = ("quest_sim_tsat_0p4s", "quest_sim_tsat_5s", "quesp_sim_b1_0p15uT")


@pytest.mark.parametrize("stem", SEQS)
def test_scanner_sequence_matches_shipped_cache(stem):
    seq = read_seq(os.path.join(PHANTOM, stem + ".seq.xz"))
    with np.load(os.path.join(PHANTOM, stem + "_events.npz")) as c:
        cache = {k: c[k] for k in c.files}
    ev, n = events_from_seq_batch(seq, int(cache["max_pulse_samples"]))
    assert n == int(cache["n_batches"])
    for field in ev._fields:
        got = np.asarray(getattr(ev, field))
        assert got.shape == cache[field].shape, field
        assert got.dtype == cache[field].dtype, field
        np.testing.assert_array_equal(got, cache[field], err_msg=field)
    assert seq.definitions["Offsets_ppm"] == pytest.approx(list(cache["offsets_ppm"]))


def test_signature_rejects_an_edited_file():
    import lzma
    with lzma.open(os.path.join(PHANTOM, "quest_sim_tsat_0p4s.seq.xz"), "rt", newline="") as f:
        text = f.read()
    edited = text.replace("This_Tsat_s 0.4", "This_Tsat_s 0.5")
    assert edited != text
    with pytest.raises(ValueError, match="md5"):
        read_seq(edited)
    read_seq(edited, check_signature=False)


def test_decode_shape_by_hand():
    # stored as is when the stored length equals num_samples
    np.testing.assert_array_equal(decode_shape([3.0, 1.0, 2.0], 3), [3.0, 1.0, 2.0])
    # 0.5 repeated: two copies plus 3 more -> five differences of 0.5, then 1.0 and -4.0
    got = decode_shape([0.0, 0.5, 0.5, 3, 1.0, -4.0], 8)
    np.testing.assert_allclose(got, [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.5, -0.5])
    # a repeat count of 0 means exactly the two copies
    np.testing.assert_array_equal(decode_shape([1.0, 1.0, 0, 2.0, 2.0, 1], 5), [1.0, 2.0, 4.0, 6.0, 8.0])
    with pytest.raises(ValueError):
        decode_shape([1.0, 1.0, 9], 5)


def test_encode_decode_round_trip():
    ramp = np.r_[np.zeros(50), np.arange(20.0), np.full(30, 19.0), [7.0]]
    packed = encode_shape(ramp)
    assert packed.size < ramp.size
    np.testing.assert_array_equal(decode_shape(packed, ramp.size), ramp)
    rng = np.random.default_rng(3)
    noise = rng.standard_normal(40)
    np.testing.assert_array_equal(encode_shape(noise), noise)


def test_block_pulse_steps_and_offsets():
    """Format 1.4 block pulses: one step of the pulse's length, split into ceil(T / 0.1 s) equal steps above
    0.1 s; delays and spoilers keep their block durations; each ADC closes a row (pulseq.py:86-115)."""
    w = SeqWriter()
    offsets = [-150.0, 0.0, 325.5]
    for f in offsets:
        w.block_pulse(42.0, 0.25, freq_offset=f)
        w.block_pulse(10.0, 0.1, freq_offset=f)
        w.spoiler(2e-3)
        w.delay(0.5)
        w.adc()
    ev, n = events_from_seq_batch(read_seq(w.text()))
    assert n == 3
    typ = [E.RF_PULSE] * 4 + [E.SPOILER, E.DELAY, E.ADC]
    for r, f in enumerate(offsets):
        np.testing.assert_array_equal(ev.type[r], typ)
        np.testing.assert_allclose(ev.dt[r], [0.25 / 3] * 3 + [0.1, 2e-3, 0.5, 0.0], rtol=1e-15)
        np.testing.assert_array_equal(ev.rf_amp[r], [42.0] * 3 + [10.0, 0, 0, 0])
        np.testing.assert_array_equal(ev.rf_freq[r], [f] * 4 + [0, 0, 0])
        np.testing.assert_array_equal(ev.adc_idx[r], [-1] * 6 + [0])


def test_phase_accumulates_from_pulse_to_pulse():
    """Each pulse's phase is its own offset less the phase the preceding pulses of the row wound up,
    dt * n * 360 * f degrees each, reduced mod 360 (pulseq.py:79-82); an ADC resets it."""
    w = SeqWriter()
    # 2.5 ms at 100 Hz winds 90 degrees; at -100 Hz, -90 = 270 mod 360. Each pulse's share is reduced,
    # the running sum is not.
    for f in (100.0, -100.0):
        w.block_pulse(5.0, 2.5e-3, freq_offset=f, phase_offset=0.25)
        w.block_pulse(5.0, 2.5e-3, freq_offset=f, phase_offset=0.25)
        w.block_pulse(5.0, 2.5e-3, freq_offset=f, phase_offset=0.0)
        w.adc()
    ev, _ = events_from_seq_batch(read_seq(w.text()))
    np.testing.assert_allclose(ev.rf_phase[0, :3], [0.25, 0.25 - math.pi / 2, -math.pi], atol=1e-15)
    np.testing.assert_allclose(ev.rf_phase[1, :3], [0.25, 0.25 - 1.5 * math.pi, -3 * math.pi], atol=1e-14)


def test_shaped_pulse_zero_samples_and_decimation():
    """Zero samples become one DELAY after the pulse; more than max_pulse_samples kept samples are
    decimated by ceil(n / max) (pulseq.py:395-422). The sample phase enters with a minus sign."""
    n_on, n_off = 1000, 120
    t = (np.arange(n_on) + 0.5) / n_on
    amp = 20.0 * np.exp(-0.5 * ((t - 0.5) / 0.2) ** 2)
    sig = np.r_[amp * np.exp(0.3j), np.zeros(n_off)]
    w = SeqWriter()
    w.shaped_pulse(sig, freq_offset=50.0)
    w.adc()
    seq = read_seq(w.text())
    ev, _ = events_from_seq_batch(seq, max_pulse_samples=300)
    stride = math.ceil(n_on / 300)                       # 4 -> 250 steps of 4 us
    n_steps = len(range(0, n_on, stride))
    assert n_steps == 250
    row = ev.type[0]
    np.testing.assert_array_equal(row, [E.RF_PULSE] * n_steps + [E.DELAY, E.ADC])
    np.testing.assert_allclose(ev.dt[0, :n_steps], stride * 1e-6, rtol=1e-12)
    np.testing.assert_allclose(ev.dt[0, n_steps], n_off * 1e-6, rtol=1e-12)
    np.testing.assert_allclose(ev.rf_amp[0, :n_steps], amp[::stride], rtol=1e-12)
    np.testing.assert_allclose(ev.rf_phase[0, :n_steps], -0.3, atol=1e-12)
    # below the limit every kept sample is a step
    ev2, _ = events_from_seq_batch(seq, max_pulse_samples=n_on)
    assert int(np.sum(ev2.type[0] == E.RF_PULSE)) == n_on


def test_single_row_form_and_padding():
    w = SeqWriter()
    for f in (0.0, 10.0):
        w.block_pulse(3.0, 0.05, freq_offset=f)
        w.adc()
    w.delay(0.2)                                         # after the last ADC: kept in the single-row form
    seq = read_seq(w.text())
    ev, n = events_from_seq(seq)
    assert n == 5 and ev.dt.dtype == np.float32
    np.testing.assert_array_equal(ev.type, [E.RF_PULSE, E.ADC, E.RF_PULSE, E.ADC, E.DELAY])
    np.testing.assert_array_equal(ev.adc_idx, [-1, 0, -1, 1, -1])
    padded, n2 = events_from_seq(seq, max_events=8)
    assert n2 == 5 and padded.type.shape == (8,)
    np.testing.assert_array_equal(padded.type[5:], E.DELAY)
    np.testing.assert_array_equal(padded.dt[5:], 0.0)
    with pytest.raises(ValueError, match="max_events"):
        events_from_seq(seq, max_events=4)
    batch, rows = events_from_seq_batch(seq)
    assert rows == 2 and batch.type.shape == (2, 2)


def test_cw_file_matches_cw_saturation():
    """The CW saturation of tutorial 03 written as a file and read back carries the same amplitude,
    offsets and total saturation time as the sequence built in code."""
    ppm = np.linspace(-4, 4, 9)
    b0, b1, tsat = 3.0, 1.0, 2.0
    code = cw_saturation(ppm, b1, tsat, b0_T=b0)
    w = SeqWriter()
    amp = float(code.rf_amp[0, 0])
    for p in ppm:
        w.block_pulse(amp, tsat, freq_offset=p * hz_per_ppm(b0))
        w.adc()
    ev, n = events_from_seq_batch(read_seq(w.text()))
    assert n == ppm.size
    rf = ev.type == E.RF_PULSE
    assert rf.sum(axis=1).tolist() == [20] * ppm.size   # 2 s in steps of 0.1 s
    np.testing.assert_allclose(np.where(rf, ev.dt, 0).sum(axis=1), tsat, rtol=1e-14)
    np.testing.assert_allclose(ev.rf_freq[:, 0], code.rf_freq[:, 0], rtol=1e-6)
    np.testing.assert_allclose(ev.rf_amp[rf], amp, rtol=1e-15)


def test_rejects_other_versions_and_bad_rows():
    text = SeqWriter().text()
    with pytest.raises(ValueError, match="1.4"):
        read_seq(text.replace("minor 4", "minor 3"), check_signature=False)
    w = SeqWriter()
    w.delay(0.01)
    bad = w.text().replace("1 1000 0 0 0 0 0 0", "1 1000 0 0 0 0 0", 1)
    with pytest.raises(ValueError, match="8 columns"):
        read_seq(bad, check_signature=False)
    with pytest.raises(ValueError, match="multiple"):
        SeqWriter().delay(1.5e-6)
    with pytest.raises(FileNotFoundError):
        read_seq("/nonexistent/file.seq")
