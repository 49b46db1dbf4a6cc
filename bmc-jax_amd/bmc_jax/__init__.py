"""BMC-JAX's simulate-and-fit path on MI355X: hand-written CDNA4 HIP kernels behind the reference API.

Subpackages mirror the reference (src/bmc_jax/__init__.py):
  simulation  SimParams, events, the BMC matrix, simulate_spectrum / simulate_voxels, the B0 spread
  sequences   cw_saturation (and pulsed_train, a built-in pulse-train builder)
  fitting     Experiment, FitSpec / FreeParam, fit, fit_batch, fit_joint, levenberg_marquardt, FitResult
  design      crlb, fisher_information, optimize_design
  io          Pulseq-CEST YAML configs (formats only)
  settings    use_x64, use_expm
The forward model, its Jacobian and the batched Levenberg-Marquardt loop run in libbmc_hip.so.
"""
from . import settings  # noqa: F401

__version__ = "0.2.0"
