"""How a sequence becomes events (src/bmc_jax/sequences/__init__.py): built in code."""
from .cache import load_events, save_events
from .saturation import (GAMMA_HZ_PER_T, GAMMA_RAD_PER_S_UT, cw_saturation, hz_per_ppm, pulsed_train,
                         scanner_train)

__all__ = ["cw_saturation", "hz_per_ppm", "pulsed_train", "scanner_train", "load_events", "save_events",
           "GAMMA_RAD_PER_S_UT", "GAMMA_HZ_PER_T"]
