"""Residual closures (src/bmc_jax/fitting/residual.py:20-52): z_sim[:, mz] - z_meas, optionally skipping
the first point. Each closure carries `.jac(x)`, the GPU kernel's forward-mode Jacobian."""
import numpy as np

from .. import _engine, _native
from ..simulation.events import slice_rows
from ..simulation.params import cast_params, n_cest_of
from ..simulation.events import cast_events
from ..simulation.voxels import tile_params
from .params import write_params


def _cast_problem(base, events, M0, dtype):
    """Downcast the whole problem (params, sequence, init state) to a dtype (residual.py:20-24)."""
    if dtype is None:
        return base, events, M0
    return cast_params(base, dtype), cast_events(events, dtype), np.asarray(M0).astype(dtype)


def spec_slots(spec, nc):
    return [_native.slot_index(s.field, s.index, nc) for s in spec.slots]


def residual_one(base, events, M0, water_z, spec, z_meas, skip_first=True):
    """r(x_scaled) -> residual vector for ONE voxel (residual.py:27-36)."""
    z_t = np.asarray(z_meas[1:] if skip_first else z_meas)
    ev = slice_rows(events, slice(1, None)) if skip_first else events
    nc = n_cest_of(base)
    slots = spec_slots(spec, nc)

    def resid(x_scaled):
        p = write_params(base, x_scaled, spec)
        return _engine.simulate_states(p, ev, M0, n_voxels=1)[0, :, water_z] - z_t

    def jac(x_scaled):
        p = write_params(base, x_scaled, spec)
        _, J = _engine.simulate_with_jac(p, ev, M0, slots, spec.scales, water_z, n_voxels=1, want_m=False)
        return J[0]

    resid.jac = jac
    return resid


def residual_batch(base, events, M0, water_z, spec, z_meas_batch, skip_first=True):
    """R(X_scaled) -> (N, n_resid) for a whole phantom; `.jac(X)` -> (N, n_resid, P) (residual.py:39-52)."""
    Z_t = np.asarray(z_meas_batch[:, 1:] if skip_first else z_meas_batch)
    n = Z_t.shape[0]
    base_b = tile_params(base, n)
    ev = slice_rows(events, slice(1, None)) if skip_first else events
    nc = n_cest_of(base)
    slots = spec_slots(spec, nc)

    def resid(X_scaled):
        p = write_params(base_b, X_scaled, spec)
        return _engine.simulate_states(p, ev, M0, n_voxels=n)[..., water_z] - Z_t

    def jac(X_scaled):
        p = write_params(base_b, X_scaled, spec)
        _, J = _engine.simulate_with_jac(p, ev, M0, slots, spec.scales, water_z, n_voxels=n, want_m=False)
        return J

    resid.jac = jac
    return resid
