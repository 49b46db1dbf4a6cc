"""The event cache: a sequence already converted to events, stored as one .npz per sequence.

Reading a Pulseq file and converting it (sequences/pulseq.py:185-428 of the reference) is slow, so the
reference ships its scanner sequences converted (examples/data/events/*.npz, described in
examples/data/README.md:88-116) and loads them with examples/_phantom.py:80-112. This module reads and
writes that format:

  type, adc_idx                         int32   [n_rows, n_events]
  dt, rf_amp, rf_phase, rf_freq         float   [n_rows, n_events]
  offsets_ppm                           float   [n_rows]   the offset of every row
  max_pulse_samples                     int     ()         the conversion's pulse sampling (optional)
  anything else                         kept as metadata (n_batches, source_seq, b1_uT, tsat_s, ...)

Files are read with numpy's default allow_pickle=False: a cache holding Python objects is refused.
"""
from __future__ import annotations

import numpy as np

from ..simulation.events import SimEvent

_FIELDS = SimEvent._fields
_INT_FIELDS = ("type", "adc_idx")


def load_events(path, keep=None, drop_reference=True, dtype=np.float32, return_meta=False):
    """Events [n_offsets, n_events] of one cached sequence.

    drop_reference: the scanner sequences acquire an unsaturated reference row first (at -300 ppm), so a
    cache holds one row per entry of `offsets_ppm` and the Z-spectrum one fewer; as the reference's
    loader does (examples/_phantom.py:105-109), the first row is dropped when the row count equals
    len(offsets_ppm).
    keep: boolean mask (or index array) over the remaining rows, applied to every field; it must be the
    same mask applied to the measurement.
    dtype: floating type of the returned fields; float32 by default, the type the reference's loader
    casts the cache to (examples/_phantom.py:107-108), so the simulated inputs are the same values.
    return_meta: also return a dict of the remaining entries (offsets_ppm after dropping and masking).
    """
    with np.load(path) as d:
        missing = [f for f in _FIELDS if f not in d.files]
        if missing:
            raise ValueError(f"{path}: not an event cache, missing {missing}")
        arrays = {f: np.asarray(d[f]) for f in _FIELDS}
        meta = {k: np.asarray(d[k]) for k in d.files if k not in _FIELDS}
    shape = arrays["type"].shape
    if len(shape) != 2 or any(a.shape != shape for a in arrays.values()):
        raise ValueError(f"{path}: event fields must share one (n_rows, n_events) shape")
    n_rows = shape[0]
    offsets = meta.get("offsets_ppm")
    if drop_reference and offsets is not None and offsets.shape == (n_rows,):
        arrays = {f: a[1:] for f, a in arrays.items()}
        offsets = offsets[1:]
        n_rows -= 1
    if keep is not None:
        keep = np.asarray(keep)
        arrays = {f: a[keep] for f, a in arrays.items()}
        if offsets is not None and offsets.shape == (n_rows,):
            offsets = offsets[keep]
    if offsets is not None:
        meta["offsets_ppm"] = offsets
    ev = SimEvent(**{f: arrays[f].astype(np.int32 if f in _INT_FIELDS else dtype) for f in _FIELDS})
    return (ev, meta) if return_meta else ev


def save_events(path, events: SimEvent, offsets_ppm, max_pulse_samples=None, **meta):
    """Write `events` ([n_rows, n_events] fields) in the cache format; offsets_ppm has one entry per row."""
    t = np.asarray(events.type)
    offsets_ppm = np.asarray(offsets_ppm, np.float64)
    if t.ndim != 2 or offsets_ppm.shape != (t.shape[0],):
        raise ValueError("events must be (n_rows, n_events) with one offset per row")
    out = {f: np.asarray(getattr(events, f)).astype(np.int32 if f in _INT_FIELDS else np.float64) for f in _FIELDS}
    out["offsets_ppm"] = offsets_ppm
    out["n_batches"] = np.asarray(t.shape[0])
    if max_pulse_samples is not None:
        out["max_pulse_samples"] = np.asarray(int(max_pulse_samples))
    for k, v in meta.items():
        out[k] = np.asarray(v)
    np.savez_compressed(path, **out)
