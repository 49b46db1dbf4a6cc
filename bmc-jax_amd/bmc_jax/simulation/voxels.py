"""The voxel batch (src/bmc_jax/simulation/voxels.py:47-151): per-voxel params and M_init, shared events.

On the device the voxel axis is simply more teams of the same launch.
"""
from typing import List

import numpy as np

from .. import _engine
from .events import SimEvent
from .params import SimParams


def stack_params(params_list: List[SimParams]) -> SimParams:
    """N per-voxel SimParams -> one with a leading voxel axis (voxels.py:47-71)."""
    if not params_list:
        raise ValueError("params_list is empty")
    return SimParams(*[np.stack([np.asarray(getattr(p, f)) for p in params_list], axis=0) for f in SimParams._fields])


def tile_params(params: SimParams, n_voxels: int) -> SimParams:
    """Replicate one tissue into a batch of n_voxels (voxels.py:74-83)."""
    return SimParams(*[np.broadcast_to(np.asarray(getattr(params, f)), (n_voxels,) + np.shape(getattr(params, f)))
                       for f in SimParams._fields])


def simulate_voxels(params_batch: SimParams, events: SimEvent, M_init_batch, return_full_trajectory: bool = False):
    """Many voxels against one sequence: (V, O, N) final states, or (V, O, T, N) (voxels.py:89-129)."""
    V = int(np.shape(params_batch.water_r1)[0])
    if return_full_trajectory:
        return _engine.simulate_trajectory(params_batch, events, M_init_batch, n_voxels=V)
    return _engine.simulate_states(params_batch, events, M_init_batch, n_voxels=V)


def extract_z_spectra(trajectory, water_z_index: int):
    """Water Mz per voxel and offset (voxels.py:135-151)."""
    return np.asarray(trajectory)[..., water_z_index]
