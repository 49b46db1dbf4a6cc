"""How a sequence becomes events (src/bmc_jax/sequences/__init__.py): built in code."""
from .saturation import cw_saturation, hz_per_ppm, pulsed_train, GAMMA_RAD_PER_S_UT, GAMMA_HZ_PER_T

__all__ = ["cw_saturation", "hz_per_ppm", "pulsed_train", "GAMMA_RAD_PER_S_UT", "GAMMA_HZ_PER_T"]
