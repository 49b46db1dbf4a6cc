"""How a sequence becomes events (src/bmc_jax/sequences/__init__.py): built in code (saturation.py) or read
from a Pulseq file (pulseq.py, which reads format 1.4 itself and also takes a pypulseq Sequence)."""
from .cache import load_events, save_events
from .pulseq import SeqFile, SeqWriter, events_from_seq, events_from_seq_batch, prep_rf_simulation, read_seq
from .saturation import (GAMMA_HZ_PER_T, GAMMA_RAD_PER_S_UT, cw_saturation, hz_per_ppm, pulsed_train,
                         scanner_train)

__all__ = ["cw_saturation", "hz_per_ppm", "pulsed_train", "scanner_train", "load_events", "save_events",
           "events_from_seq", "events_from_seq_batch", "prep_rf_simulation", "read_seq", "SeqFile", "SeqWriter",
           "GAMMA_RAD_PER_S_UT", "GAMMA_HZ_PER_T"]
