"""Pulseq sequences as events: a reader of Pulseq files (format 1.4) that needs no pypulseq, and the
conversion of a sequence into solver events (src/bmc_jax/sequences/pulseq.py:20-428).

The conversion walks the blocks in order (pulseq.py:59-155, 224-326):
  * a block with an ADC becomes an ADC event of zero duration; it resets the accumulated RF phase, and in
    the batch form it closes the row, so every ADC ends one offset;
  * a block with an RF pulse becomes RF_PULSE events: a shaped pulse one event per kept sample, a block
    pulse one event (split into equal steps of at most 0.1 s), followed by a DELAY for its zero samples;
  * a block with a z gradient becomes one SPOILER event of the block's duration;
  * any other block becomes a DELAY of the block's duration.
The checks go in that order: a block with an ADC and an RF pulse counts as an ADC.

The converter reads a sequence through the attributes pypulseq's `Sequence` offers (`block_events`,
`get_block(i)` with `.adc`, `.rf`, `.gz`, `.block_duration`, and on the RF `.signal`, `.t`,
`.freq_offset`, `.phase_offset`), so it takes a pypulseq `Sequence` where that library is installed and
the `SeqFile` of `read_seq` everywhere.
"""
from __future__ import annotations

import hashlib
import lzma
import math
import os
from types import SimpleNamespace

import numpy as np

from ..simulation.events import EventType, SimEvent

MAX_BLOCK_STEP_S = 0.1       # block pulses longer than this are split into equal steps (pulseq.py:86-101)
ZERO_SAMPLE = 1e-6           # |signal| at or below this is a zero sample (pulseq.py:395-403)


# ------------------------------------------------------------------------------------------------
# Reading a file
# ------------------------------------------------------------------------------------------------

def decode_shape(packed, num_samples):
    """One shape's samples from its stored form (Pulseq 1.4 shape compression).

    A shape whose stored length equals num_samples is stored as is. Otherwise it is stored as the run-length
    encoded first difference of the samples: a value that appears twice in a row is followed by the count
    of further copies of it, and the samples are the running sum of the expanded differences.
    """
    packed = np.asarray(packed, np.float64)
    n = int(num_samples)
    if packed.size == n:
        return packed.copy()
    diff = np.empty(n, np.float64)
    i = j = 0
    m = packed.size
    vals = packed.tolist()
    while i < m:
        v = vals[i]
        if i + 2 < m and vals[i + 1] == v:
            reps = 2 + int(round(vals[i + 2]))
            if j + reps > n:
                raise ValueError(f"shape data expands past its {n} samples")
            diff[j:j + reps] = v
            j += reps
            i += 3
        else:
            if j >= n:
                raise ValueError(f"shape data expands past its {n} samples")
            diff[j] = v
            j += 1
            i += 1
    if j != n:
        raise ValueError(f"shape data expands to {j} samples, not the {n} it declares")
    return np.cumsum(diff)


def _numbers(tokens):
    out = []
    for t in tokens:
        try:
            out.append(float(t))
        except ValueError:
            out.append(t)
    return out


class SeqFile:
    """A Pulseq file (format 1.4), read into its event tables.

    blocks       {block id: (duration in raster units, rf, gx, gy, gz, adc, ext)} event ids, 0 = none
    rf           {id: (amplitude Hz, mag shape, phase shape, time shape, delay us, freq Hz, phase rad)}
    trap         {id: (amplitude Hz/m, rise us, flat us, fall us, delay us)}
    grad         {id: (amplitude Hz/m, shape, time shape, delay us)}
    adc          {id: (samples, dwell ns, delay us, freq Hz, phase rad)}
    shapes       {id: decoded samples}
    definitions  {name: value or list of values}
    """

    def __init__(self, text, check_signature=True):
        self.version = None
        self.definitions = {}
        self.blocks = {}
        self.rf, self.trap, self.grad, self.adc, self.shapes = {}, {}, {}, {}, {}
        self.signature = None
        self._signals = {}
        self._parse(text)
        if self.version is None or self.version[:2] != (1, 4):
            raise ValueError(f"reads Pulseq format 1.4 files; this one is version {self.version}")
        if check_signature and self.signature is not None:
            kind, value = self.signature
            if kind.lower() != "md5":
                raise ValueError(f"unknown signature type {kind!r}")
            cut = text.find("\n[SIGNATURE]")
            if hashlib.md5(text[:cut].encode()).hexdigest() != value.lower():
                raise ValueError("the file does not match its md5 signature: it was changed after writing")
        self.block_duration_raster = float(self._definition("BlockDurationRaster", 1e-5))
        self.rf_raster_time = float(self._definition("RadiofrequencyRasterTime", 1e-6))
        self.grad_raster_time = float(self._definition("GradientRasterTime", 1e-5))
        self.adc_raster_time = float(self._definition("AdcRasterTime", 1e-7))

    def _definition(self, name, default):
        v = self.definitions.get(name, default)
        return v[0] if isinstance(v, list) else v

    def _parse(self, text):
        section = None
        shape_id = shape_n = None
        shape_vals = []

        def close_shape():
            if shape_id is not None:
                self.shapes[shape_id] = decode_shape(shape_vals, shape_n)

        rows = {"RF": (self.rf, 8), "TRAP": (self.trap, 6), "GRADIENTS": (self.grad, 5), "ADC": (self.adc, 6)}
        version = {}
        for lineno, raw in enumerate(text.splitlines(), 1):
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            if line.startswith("[") and line.endswith("]"):
                if section == "SHAPES":
                    close_shape()
                    shape_id = None
                section = line[1:-1].upper()
                continue
            tok = line.split()
            if section == "VERSION":
                version[tok[0]] = int(tok[1])
            elif section == "DEFINITIONS":
                vals = _numbers(tok[1:])
                self.definitions[tok[0]] = vals[0] if len(vals) == 1 else vals
            elif section == "BLOCKS":
                if len(tok) != 8:
                    raise ValueError(f"line {lineno}: a 1.4 block has 8 columns (id, duration, rf, gx, gy, gz, "
                                     f"adc, ext), this one {len(tok)}")
                self.blocks[int(tok[0])] = tuple(int(t) for t in tok[1:])
            elif section in rows:
                table, width = rows[section]
                if len(tok) != width:
                    raise ValueError(f"line {lineno}: a [{section}] row has {width} columns, this one {len(tok)}")
                table[int(tok[0])] = tuple(float(t) for t in tok[1:])
            elif section == "SHAPES":
                if tok[0] == "shape_id":
                    close_shape()
                    shape_id, shape_n, shape_vals = int(tok[1]), None, []
                elif tok[0] == "num_samples":
                    shape_n = int(tok[1])
                else:
                    shape_vals.extend(float(t) for t in tok)
            elif section == "SIGNATURE":
                if tok[0].lower() == "type":
                    self.signature = (tok[1], None)
                elif tok[0].lower() == "hash":
                    self.signature = ((self.signature or ("md5",))[0], tok[1])
            elif section == "DELAYS":
                raise ValueError("a [DELAYS] section belongs to Pulseq files before format 1.4")
            # [EXTENSIONS] (labels, triggers) does not enter the simulation and is skipped
        if section == "SHAPES":
            close_shape()
        if version:
            self.version = (version.get("major"), version.get("minor"), version.get("revision"))

    # --- the block view a converter reads (the attributes pypulseq's Sequence offers) ---------------

    @property
    def block_events(self):
        return self.blocks

    def _shape(self, sid, what):
        if sid not in self.shapes:
            raise ValueError(f"{what} names shape {sid}, which the file does not hold")
        return self.shapes[sid]

    def rf_event(self, rf_id):
        """The RF pulse `rf_id`: complex signal (Hz), sample times (s), delay (s), freq and phase offsets.

        Pulses that share amplitude and shapes (a train played at many offsets, or with phase cycling)
        share one signal array, read-only; `shape_key` names that array for a converter's memo.
        """
        amp, mag_id, ph_id, time_id, delay_us, freq, phase = self.rf[rf_id]
        key = (amp, int(mag_id), int(ph_id), int(time_id))
        if key not in self._signals:
            mag = self._shape(key[1], f"RF {rf_id}")
            ph = self._shape(key[2], f"RF {rf_id}")
            if mag.size != ph.size:
                raise ValueError(f"RF {rf_id}: magnitude and phase shapes differ in length ({mag.size}, "
                                 f"{ph.size})")
            signal = amp * mag * np.exp(1j * 2 * np.pi * ph)
            if key[3] == 0:
                t = (np.arange(1, mag.size + 1) - 0.5) * self.rf_raster_time
            else:
                t = self._shape(key[3], f"RF {rf_id}") * self.rf_raster_time
            signal.flags.writeable = False
            t.flags.writeable = False
            self._signals[key] = (signal, t)
        signal, t = self._signals[key]
        return SimpleNamespace(type="rf", signal=signal, t=t, delay=delay_us * 1e-6, freq_offset=freq,
                               phase_offset=phase, shape_key=(id(self), key))

    def get_block(self, block_id):
        """Block `block_id` (1-based) with its events; an absent event is None."""
        dur, rf, gx, gy, gz, adc, ext = self.blocks[block_id]

        def grad(gid):
            if gid == 0:
                return None
            if gid in self.trap:
                a, rise, flat, fall, delay = self.trap[gid]
                return SimpleNamespace(type="trap", amplitude=a, rise_time=rise * 1e-6, flat_time=flat * 1e-6,
                                       fall_time=fall * 1e-6, delay=delay * 1e-6)
            if gid in self.grad:
                a, sid, tid, delay = self.grad[gid]
                return SimpleNamespace(type="grad", amplitude=a, shape_id=int(sid), time_id=int(tid),
                                       delay=delay * 1e-6)
            raise ValueError(f"block {block_id} names gradient {gid}, which the file does not hold")

        adc_ev = None
        if adc:
            n, dwell, delay, freq, phase = self.adc[adc]
            adc_ev = SimpleNamespace(type="adc", num_samples=int(n), dwell=dwell * 1e-9, delay=delay * 1e-6,
                                     freq_offset=freq, phase_offset=phase)
        return SimpleNamespace(block_duration=dur * self.block_duration_raster,
                               rf=self.rf_event(rf) if rf else None, gx=grad(gx), gy=grad(gy), gz=grad(gz),
                               adc=adc_ev)


def read_seq(source, check_signature=True) -> SeqFile:
    """A Pulseq 1.4 file from a path (plain, or xz-compressed when the name ends in .xz) or from its text.

    check_signature: compare the file with the md5 hash of its [SIGNATURE] section, when it has one.
    """
    if isinstance(source, (str, os.PathLike)) and os.path.exists(source):
        path = os.fspath(source)
        opener = lzma.open if path.endswith(".xz") else open
        with opener(path, "rt", newline="") as f:
            text = f.read()
    elif isinstance(source, str) and "[BLOCKS]" in source:
        text = source
    else:
        raise FileNotFoundError(f"no Pulseq file at {source!r}")
    return SeqFile(text, check_signature=check_signature)


# ------------------------------------------------------------------------------------------------
# Writing a file
# ------------------------------------------------------------------------------------------------

def encode_shape(samples):
    """The stored form of a shape (the inverse of decode_shape): the first difference, run-length encoded,
    when that is shorter than the samples; the samples themselves otherwise."""
    x = np.asarray(samples, np.float64)
    d = np.diff(x, prepend=0.0).tolist()
    out, i = [], 0
    while i < len(d):
        j = i
        while j + 1 < len(d) and d[j + 1] == d[i]:
            j += 1
        run = j - i + 1
        out.extend([d[i], d[i], float(run - 2)] if run >= 2 else [d[i]])
        i = j + 1
    return np.asarray(out) if len(out) < x.size else x.copy()


def _fmt(v):
    v = float(v)
    return str(int(v)) if v.is_integer() and abs(v) < 1e15 else repr(v)


class SeqWriter:
    """Builds a Pulseq 1.4 file block by block: block pulses, shaped pulses, z spoilers, ADCs and delays.

    Durations are rounded to the block raster (an error if they are off it by more than 1e-6 of a raster);
    shaped pulses are sampled on the RF raster. Identical shapes and events are stored once.
    """

    def __init__(self, rf_raster=1e-6, block_raster=1e-5, grad_raster=1e-5, adc_raster=1e-7, definitions=None):
        self.rf_raster, self.block_raster = float(rf_raster), float(block_raster)
        self.grad_raster, self.adc_raster = float(grad_raster), float(adc_raster)
        self.definitions = dict(definitions or {})
        self._blocks, self._shapes, self._rf, self._trap, self._adc = [], {}, {}, {}, {}

    def _units(self, seconds, raster, what):
        k = round(float(seconds) / raster)
        if abs(k * raster - float(seconds)) > 1e-6 * raster:
            raise ValueError(f"{what} {seconds} s is not a multiple of the {raster} s raster")
        return int(k)

    def _id(self, table, row):
        return table.setdefault(row, len(table) + 1)

    def _shape(self, samples):
        return self._id(self._shapes, tuple(np.asarray(samples, np.float64).tolist()))

    def _block(self, dur_units, rf=0, gz=0, adc=0):
        self._blocks.append((dur_units, rf, 0, 0, gz, adc, 0))

    def delay(self, seconds):
        self._block(self._units(seconds, self.block_raster, "delay"))

    def block_pulse(self, amp_hz, duration, freq_offset=0.0, phase_offset=0.0):
        """A constant pulse of amp_hz (B1 in Hz) for `duration`, stored in the two-sample form of format 1.4."""
        n = self._units(duration, self.rf_raster, "block pulse")
        rf = self._id(self._rf, (float(amp_hz), self._shape([1.0, 1.0]), self._shape([0.0, 0.0]),
                                 self._shape([0.0, float(n)]), 0.0, float(freq_offset), float(phase_offset)))
        self._block(int(math.ceil(n * self.rf_raster / self.block_raster - 1e-9)), rf=rf)

    def shaped_pulse(self, signal_hz, freq_offset=0.0, phase_offset=0.0):
        """A pulse with one complex sample (Hz) per RF raster step."""
        sig = np.asarray(signal_hz, np.complex128)
        amp = float(np.max(np.abs(sig)))
        if sig.size < 2 or amp == 0.0:
            raise ValueError("a shaped pulse needs two samples or more and a nonzero amplitude")
        mag = np.abs(sig) / amp
        ph = np.mod(np.angle(sig) / (2 * np.pi), 1.0)
        rf = self._id(self._rf, (amp, self._shape(mag), self._shape(ph), 0, 0.0, float(freq_offset),
                                 float(phase_offset)))
        self._block(int(math.ceil(sig.size * self.rf_raster / self.block_raster - 1e-9)), rf=rf)

    def spoiler(self, duration, amplitude_hz_per_m=1e6):
        """A trapezoid on z filling `duration`: ramps of a quarter of it each, on the gradient raster."""
        n = self._units(duration, self.grad_raster, "spoiler")
        ramp = max(1, n // 4)
        us = lambda k: k * self.grad_raster * 1e6
        gz = self._id(self._trap, (float(amplitude_hz_per_m), us(ramp), us(n - 2 * ramp), us(ramp), 0.0))
        self._block(self._units(duration, self.block_raster, "spoiler"), gz=gz)

    def adc(self, num_samples=1, dwell=1e-5, delay=1e-5):
        """A readout of num_samples samples; its block lasts delay + num_samples * dwell (on the block raster)."""
        adc = self._id(self._adc, (int(num_samples), float(dwell) * 1e9, float(delay) * 1e6, 0.0, 0.0))
        total = float(delay) + int(num_samples) * float(dwell)
        self._block(int(math.ceil(total / self.block_raster - 1e-9)), adc=adc)

    def text(self):
        """The file's text, ending with its md5 signature."""
        L = ["# Pulseq sequence file", "# Written by bmc_jax.sequences.pulseq.SeqWriter", "",
             "[VERSION]", "major 1", "minor 4", "revision 2", "", "[DEFINITIONS]"]
        defs = {"AdcRasterTime": self.adc_raster, "BlockDurationRaster": self.block_raster,
                "GradientRasterTime": self.grad_raster, "RadiofrequencyRasterTime": self.rf_raster}
        defs.update(self.definitions)
        for k in sorted(defs):
            v = defs[k]
            vals = v if isinstance(v, (list, tuple, np.ndarray)) else [v]
            L.append(k + " " + " ".join(x if isinstance(x, str) else _fmt(x) for x in vals))
        L += ["", "# Format of blocks:", "# NUM DUR RF  GX  GY  GZ  ADC  EXT", "[BLOCKS]"]
        L += [" ".join(str(int(x)) for x in (i + 1,) + b) for i, b in enumerate(self._blocks)]
        tables = [("RF", "# id amplitude mag_id phase_id time_shape_id delay freq phase", self._rf),
                  ("TRAP", "# id amplitude rise flat fall delay", self._trap),
                  ("ADC", "# id num dwell delay freq phase", self._adc)]
        for name, head, table in tables:
            if table:
                L += ["", head, f"[{name}]"]
                L += [" ".join([str(i)] + [_fmt(x) for x in row]) for row, i in sorted(table.items(),
                                                                                       key=lambda kv: kv[1])]
        if self._shapes:
            L += ["", "[SHAPES]"]
            for samples, i in sorted(self._shapes.items(), key=lambda kv: kv[1]):
                L += ["", f"shape_id {i}", f"num_samples {len(samples)}"]
                L += [_fmt(x) for x in encode_shape(samples)]
        body = "\n".join(L) + "\n"
        digest = hashlib.md5(body.encode()).hexdigest()
        return body + "\n[SIGNATURE]\nType md5\nHash " + digest + "\n"

    def write(self, path):
        with open(path, "w", newline="") as f:
            f.write(self.text())


# ------------------------------------------------------------------------------------------------
# Sequence -> events
# ------------------------------------------------------------------------------------------------

def prep_rf_simulation(block, max_pulse_samples: int):
    """(amp, phase, dt, delay_after) of the block's RF pulse as the solver steps it (pulseq.py:367-428).

    Zero samples (|signal| <= 1e-6) are dropped and their time is returned as delay_after. A pulse whose
    kept samples all share one amplitude and one phase is a block pulse: one step, of the raster spacing
    when it has two samples (how format 1.4 stores a block pulse) or of the kept samples' whole length
    otherwise. A shaped pulse longer than max_pulse_samples keeps every ceil(n / max_pulse_samples)-th
    sample, each step that many raster spacings long.
    """
    sig = np.asarray(block.rf.signal)
    amp = np.abs(sig)
    ph = np.angle(sig)
    t = np.asarray(block.rf.t, np.float64)
    spacing = t[1] - t[0]
    keep = amp > ZERO_SAMPLE
    n_kept = int(np.count_nonzero(keep))
    delay_after = (amp.size - n_kept) * spacing
    amp, ph = amp[keep], ph[keep]
    distinct = max(np.unique(amp).size, np.unique(ph).size)
    if distinct == 1:
        return amp[0], ph[0], (spacing if amp.size == 2 else spacing * amp.size), delay_after
    if amp.size > max_pulse_samples:
        stride = int(math.ceil(amp.size / max_pulse_samples))
        return amp[::stride], ph[::stride], spacing * stride, delay_after
    return amp, ph, spacing, delay_after


def _rows(seq, max_pulse_samples, split_at_adc):
    """Walk the blocks; one list of (type, dt, amp, phase, freq, adc_idx) per row (pulseq.py:59-155)."""
    rows, row = [], []
    adc_count = 0
    accum = 0.0
    n_blocks = len(seq.block_events)
    memo = {}
    for b in range(1, n_blocks + 1):
        block = seq.get_block(b)
        if block.adc is not None:
            row.append((EventType.ADC, 0.0, 0.0, 0.0, 0.0, adc_count))
            accum = 0.0
            if split_at_adc:
                rows.append(row)
                row = []
            else:
                adc_count += 1
        elif block.rf is not None:
            key = getattr(block.rf, "shape_key", None)
            if key is None:
                amp, ph, dt, delay_after = prep_rf_simulation(block, max_pulse_samples)
            else:
                if key not in memo:
                    memo[key] = prep_rf_simulation(block, max_pulse_samples)
                amp, ph, dt, delay_after = memo[key]
            freq = float(block.rf.freq_offset)
            offset, before = float(block.rf.phase_offset), accum
            n = np.size(amp)
            accum += ((float(dt) * n * 360 * freq) % 360) / 180 * np.pi
            if n == 1:
                total = float(dt)
                steps = int(math.ceil(total / MAX_BLOCK_STEP_S)) if total > MAX_BLOCK_STEP_S else 1
                a, p = float(np.squeeze(amp)), -float(np.squeeze(ph)) + offset - before
                row.extend([(EventType.RF_PULSE, total / steps if steps > 1 else total, a, p, freq, -1)] * steps)
            else:
                row.extend((EventType.RF_PULSE, float(dt), float(a), -float(p) + offset - before, freq, -1)
                           for a, p in zip(amp, ph))
            if delay_after > 1e-12:
                row.append((EventType.DELAY, float(delay_after), 0.0, 0.0, 0.0, -1))
        elif block.gz is not None:
            row.append((EventType.SPOILER, float(block.block_duration), 0.0, 0.0, 0.0, -1))
        else:
            row.append((EventType.DELAY, float(block.block_duration), 0.0, 0.0, 0.0, -1))
    if not split_at_adc:
        rows.append(row)
    return rows


def _as_events(rows, width, float_dtype):
    n = len(rows)
    typ = np.zeros((n, width), np.int32)
    adc = np.full((n, width), -1, np.int32)
    flt = np.zeros((4, n, width), np.float64)
    for i, r in enumerate(rows):
        if not r:
            continue
        a = np.asarray(r, np.float64)
        typ[i, :len(r)] = a[:, 0]
        flt[:, i, :len(r)] = a[:, 1:5].T
        adc[i, :len(r)] = a[:, 5]
    flt = flt.astype(float_dtype)
    return SimEvent(type=typ, dt=flt[0], rf_amp=flt[1], rf_phase=flt[2], rf_freq=flt[3], adc_idx=adc)


def _sequence(seq):
    return seq if hasattr(seq, "get_block") else read_seq(seq)


def events_from_seq(seq, max_pulse_samples: int = 500, max_events=None, dtype=np.float32):
    """(events, n_events): the whole sequence as one row of events (pulseq.py:20-183).

    seq: a SeqFile, a pypulseq Sequence, or a path to a .seq file. ADC events count up (adc_idx 0, 1, ...).
    max_events pads the row with zero-duration DELAYs to that length (ValueError if the sequence has more).
    The fields are 1-D (n_events,), floats in `dtype` (float32 as the reference's converter returns them).
    """
    (row,) = _rows(_sequence(seq), max_pulse_samples, split_at_adc=False)
    n = len(row)
    width = n
    if max_events is not None:
        if n > int(max_events):
            raise ValueError(f"the sequence converts to {n} events, more than max_events={max_events}")
        width = int(max_events)
    ev = _as_events([row], width, dtype)
    return SimEvent(*[np.asarray(f)[0] for f in ev]), n


def events_from_seq_batch(seq, max_pulse_samples: int = 500):
    """(events, n_rows): one row per ADC, the events since the previous ADC up to and including it
    (pulseq.py:185-365); blocks after the last ADC belong to no row. Rows are padded at the end with
    zero-duration DELAYs to the longest; fields (n_rows, width), floats in float64, adc_idx 0 on each ADC.
    """
    rows = _rows(_sequence(seq), max_pulse_samples, split_at_adc=True)
    width = max((len(r) for r in rows), default=0)
    return _as_events(rows, width, np.float64), len(rows)
