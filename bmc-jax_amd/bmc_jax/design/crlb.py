"""The Cramer-Rao lower bound (src/bmc_jax/design/crlb.py:30-120): F = J^T J / sigma^2 with J from the GPU
kernel's forward tangents, inverted in float64 on the host."""
from dataclasses import dataclass

import numpy as np

from .. import _engine
from ..fitting.params import write_params
from ..fitting.residual import spec_slots
from ..simulation.params import n_cest_of
from ..simulation.simulate import simulate_spectrum


def _require_one_plain_sample(experiment, spec):
    """The bound is defined here for a single sample whose values are plain per-sample values
    (crlb.py:30-35)."""
    n_dims = np.ndim(experiment.params.water_r1)
    if n_dims:
        raise ValueError(f"crlb needs the parameters of one sample (scalar fields); these have {n_dims} batch "
                         "axis")
    if not spec.independent:
        raise ValueError("crlb takes plain values per sample only: no shared or per-block scope, no samples=, "
                         "no factor or offset mode, no m_start or b0_spread")


def model_of(experiment, spec):
    """The model x (physical) -> Z at the experiment's offsets (crlb.py:38-47); the returned function
    carries .jac(x) = dZ/dx from the GPU kernel's tangents."""
    _require_one_plain_sample(experiment, spec)
    base, events, m0, mz = experiment.params, experiment.events, experiment.m0, experiment.mz_loc
    scales = spec.scales
    slots = spec_slots(spec, n_cest_of(base))

    def placed(x):
        return write_params(base, np.asarray(x, np.float64) / scales, spec)

    def z(x):
        return simulate_spectrum(placed(x), events, m0, return_full_trajectory=False)[:, mz]

    def jac(x):
        _, J = _engine.simulate_with_jac(placed(x), events, m0, slots, np.ones(len(slots)), mz, n_voxels=1,
                                         want_m=False)
        return J[0]

    z.jac = jac
    return z


def fisher_information(experiment, spec, x, sigma):
    """(F, J): F = J^T J / sigma^2 at x (physical units), J = dZ/dx (crlb.py:50-58)."""
    J = np.asarray(model_of(experiment, spec).jac(np.asarray(x, np.float64)), np.float64)
    return (J.T @ J) / float(sigma) ** 2, J


@dataclass(frozen=True)
class CRLB:
    """The bound at the values x (physical units): sd is the smallest standard deviation an unbiased
    estimate of each value can have; corr the correlations of those estimates; fisher and jacobian the
    matrices it came from (crlb.py:61-85)."""
    names: tuple
    x: np.ndarray
    sd: np.ndarray
    corr: np.ndarray
    fisher: np.ndarray
    jacobian: np.ndarray

    @property
    def rel_sd(self):
        """sd relative to the value, sd / |x|."""
        return self.sd / np.abs(self.x)

    def __getitem__(self, name):
        """The bound of one value, by its FreeParam name."""
        return float(self.sd[list(self.names).index(name)])

    def summary(self):
        """One line per value: its name, the value, the bound and the bound in percent of the value."""
        lines = ["parameter".rjust(10) + " " + "value".rjust(12) + " " + "CRLB (SD)".rjust(12) + " "
                 + "relative".rjust(9)]
        for name, value, bound, rel in zip(self.names, self.x, self.sd, self.rel_sd):
            lines.append(f"{name:>10} {value:>12.6g} {bound:>12.4g} {100.0 * rel:>8.2f}%")
        return "\n".join(lines)


def crlb(experiment, spec, sigma, x=None):
    """The Cramer-Rao bound of spec's values for one sample at x (default: each value's x0), with noise
    sigma on every point (crlb.py:88-120)."""
    at = np.asarray(spec.x0_phys if x is None else x, float)
    F, J = fisher_information(experiment, spec, at, sigma)
    F = np.asarray(F, np.float64)
    cov = np.linalg.inv(F)
    sd = np.sqrt(np.maximum(np.diag(cov), 0.0))
    corr = cov / sd[:, None] / sd[None, :]
    return CRLB(tuple(s.name for s in spec.slots), at, sd, corr, F, np.asarray(J))
