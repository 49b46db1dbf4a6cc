"""The Cramer-Rao lower bound (src/bmc_jax/design/crlb.py:30-120): F = J^T J / sigma^2 with J from the GPU
kernel's forward tangents, inverted in float64 on the host."""
from dataclasses import dataclass

import numpy as np

from .. import _engine
from ..fitting.params import write_params
from ..fitting.residual import spec_slots
from ..simulation.params import n_cest_of
from ..simulation.simulate import simulate_spectrum


def _check(experiment, spec):
    if np.ndim(experiment.params.water_r1) != 0:
        raise ValueError("the bound is computed for one sample: give an Experiment whose params carry one sample")
    if not spec.independent:
        raise ValueError("the bound supports plain values per sample (FreeParam scope='sample', mode='value', no "
                         "samples=, no m_start or b0_spread)")


def model_of(experiment, spec):
    """x (physical) -> Z (crlb.py:38-47); carries .jac(x) = dZ/dx from the GPU."""
    _check(experiment, spec)
    base, events, m0, mz = experiment.params, experiment.events, experiment.m0, experiment.mz_loc
    scales = spec.scales
    slots = spec_slots(spec, n_cest_of(base))

    def z(x):
        p = write_params(base, np.asarray(x, np.float64) / scales, spec)
        return simulate_spectrum(p, events, m0, return_full_trajectory=False)[:, mz]

    def jac(x):
        p = write_params(base, np.asarray(x, np.float64) / scales, spec)
        _, J = _engine.simulate_with_jac(p, events, m0, slots, [1.0] * len(slots), mz, n_voxels=1, want_m=False)
        return J[0]

    z.jac = jac
    return z


def fisher_information(experiment, spec, x, sigma):
    """F = J^T J / sigma^2 at x (physical) and J (crlb.py:50-58)."""
    J = np.asarray(model_of(experiment, spec).jac(np.asarray(x, np.float64)), np.float64)
    return J.T @ J / sigma ** 2, J


@dataclass(frozen=True)
class CRLB:
    names: tuple
    x: np.ndarray
    sd: np.ndarray
    corr: np.ndarray
    fisher: np.ndarray
    jacobian: np.ndarray

    @property
    def rel_sd(self):
        return self.sd / np.abs(self.x)

    def __getitem__(self, name):
        return float(self.sd[self.names.index(name)])

    def summary(self):
        rows = [f"{'parameter':>10} {'value':>12} {'CRLB (SD)':>12} {'relative':>9}"]
        for n, x, s, r in zip(self.names, self.x, self.sd, self.rel_sd):
            rows.append(f"{n:>10} {x:>12.6g} {s:>12.4g} {100 * r:>8.2f}%")
        return "\n".join(rows)


def crlb(experiment, spec, sigma, x=None):
    """The bound of spec's values for one sample (crlb.py:88-120)."""
    x = np.asarray(spec.x0_phys if x is None else x, float)
    F, J = fisher_information(experiment, spec, x, sigma)
    F = np.asarray(F, np.float64)
    cov = np.linalg.inv(F)
    sd = np.sqrt(np.clip(np.diag(cov), 0.0, None))
    corr = cov / np.outer(sd, sd)
    return CRLB(tuple(s.name for s in spec.slots), x, sd, corr, F, np.asarray(J))
