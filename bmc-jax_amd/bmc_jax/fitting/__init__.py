"""Fit selected parameters of an Experiment to measured Z-spectra (src/bmc_jax/fitting/__init__.py).

lm.py imports nothing from the package (the reference enforces this in tests/test_structure.py).
"""
from .experiment import Experiment
from .params import FitSpec, FreeParam, write_params
from .residual import residual_one, residual_batch
from .lm import levenberg_marquardt
from .independent import fit, fit_one, fit_batch, fit_batch_sharded, fit_batch_adam
from .joint import fit_joint
from .result import FitResult

__all__ = [
    "Experiment",
    "FreeParam", "FitSpec", "write_params",
    "residual_one", "residual_batch",
    "levenberg_marquardt",
    "fit", "FitResult",
    "fit_joint",
    "fit_one", "fit_batch", "fit_batch_sharded", "fit_batch_adam",
]
