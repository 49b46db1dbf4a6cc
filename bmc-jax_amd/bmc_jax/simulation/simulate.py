"""The event scan (src/bmc_jax/simulation/simulate.py:22-148), run by the HIP kernel.

Semantics kept exactly: an event evolves the state only if dt > 1e-9; GRADIENT and SPOILER zero the
transverse components of water and the CEST pools (the MT pool untouched); last_M is the state after
every event; an ADC with reset_after_adc sets the running state back to M_init for the next event.
"""
from typing import NamedTuple

import numpy as np

from .. import _engine, settings
from .events import SimEvent
from .params import SimParams


class SimState(NamedTuple):
    """State carried through the scan (simulate.py:17-20)."""
    M: np.ndarray
    last_M: np.ndarray


def _check_single(params):
    if np.ndim(params.water_r1) != 0:
        raise ValueError("simulate_spectrum takes the parameters of one sample; use simulate_voxels for a batch")


def _plain_guard():
    if settings.expm_backend() == "plain" and settings.x64_enabled():
        raise ValueError('use_expm("plain") is float32 only (x64 is on)')


def simulate_sequence(params: SimParams, events: SimEvent, M_init, reset_after_adc: bool = True,
                      return_full_trajectory: bool = True):
    """ONE offset row (1-D event fields): trajectory (n_events, N) or final last_M (N,) (simulate.py:80-132)."""
    _check_single(params)
    _plain_guard()
    row = SimEvent(*[np.asarray(f)[None, :] for f in events])
    if return_full_trajectory:
        return _engine.simulate_trajectory(params, row, M_init, n_voxels=1, reset_after_adc=reset_after_adc)[0, 0]
    return _engine.simulate_states(params, row, M_init, n_voxels=1, reset_after_adc=reset_after_adc)[0, 0]


def simulate_spectrum(params: SimParams, events: SimEvent, M_init, return_full_trajectory: bool = True):
    """Every offset row (simulate.py:134-148): (O, T, N) trajectories or (O, N) final states."""
    _check_single(params)
    _plain_guard()
    if return_full_trajectory:
        return _engine.simulate_trajectory(params, events, M_init, n_voxels=1)[0]
    return _engine.simulate_states(params, events, M_init, n_voxels=1)[0]
