"""Writes the phantom fixtures from the reference's shipped measurement (run where the reference is mounted):

    python tests/golden/make_phantom.py [/path/to/BMC-JAX]

tests/golden/phantom/phantom_7tubes.npz   the 2026-09-27 seven-tube phantom scan (examples/data/phantom_7tubes.npz):
                                          averaged Z-spectra (QUESP and QUEST ladders), per-tube constants
tests/golden/phantom/<stem>_events.npz    the 14 scanner sequences converted to events (examples/data/events/)
tests/golden/phantom/<stem>.seq.xz        three of those sequences as their Pulseq files (examples/data/seqs/),
                                          xz-compressed byte for byte: the inputs the event caches were made from

Every file is read with numpy's allow_pickle=False (plain arrays only) and written back with the same keys
and dtypes, compressed. Only data moves; nothing of the reference's code does.
"""
import glob
import lzma
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "phantom")
SEQS = ("quest_sim_tsat_0p4s", "quest_sim_tsat_5s", "quesp_sim_b1_0p15uT")


def copy_arrays(src, dst):
    with np.load(src, allow_pickle=False) as d:
        arrays = {k: np.asarray(d[k]) for k in d.files}
    np.savez_compressed(dst, **arrays)
    return sorted(arrays)


def main(ref="/root/reference"):
    data = os.path.join(ref, "examples", "data")
    os.makedirs(OUT, exist_ok=True)
    keys = copy_arrays(os.path.join(data, "phantom_7tubes.npz"), os.path.join(OUT, "phantom_7tubes.npz"))
    print("phantom:", keys)
    for f in sorted(glob.glob(os.path.join(data, "events", "*_events.npz"))):
        copy_arrays(f, os.path.join(OUT, os.path.basename(f)))
        print("events:", os.path.basename(f))
    for stem in SEQS:
        with open(os.path.join(data, "seqs", stem + ".seq"), "rb") as f:
            raw = f.read()
        with lzma.open(os.path.join(OUT, stem + ".seq.xz"), "wb", preset=9) as f:
            f.write(raw)
        print("sequence:", stem + ".seq")


if __name__ == "__main__":
    main(*sys.argv[1:])
