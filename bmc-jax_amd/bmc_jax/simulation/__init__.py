"""The forward model: SimParams, events, the BMC matrix, the event scan, voxel batches, the B0 spread.

Mirrors src/bmc_jax/simulation/__init__.py of the reference; the scan and its Jacobian run in the HIP
kernels of libbmc_hip.so.
"""
from .params import SimParams, to_sim_params, initial_state, cast_params
from .events import SimEvent, EventType, pad_events, cast_events, repeat_events
from .matrix import bmc_matrix, lorentzian, super_lorentzian
from .expm import propagate, expm_library, expm_plain
from .simulate import simulate_sequence, simulate_spectrum
from .voxels import simulate_voxels, tile_params, stack_params, extract_z_spectra
from .spread import B0Spread, simulate_spectrum_spread

__all__ = ["SimParams", "to_sim_params", "initial_state", "cast_params",
           "SimEvent", "EventType", "pad_events", "cast_events", "repeat_events",
           "bmc_matrix", "lorentzian", "super_lorentzian",
           "propagate", "expm_library", "expm_plain",
           "simulate_sequence", "simulate_spectrum",
           "simulate_voxels", "tile_params", "stack_params", "extract_z_spectra",
           "B0Spread", "simulate_spectrum_spread"]
