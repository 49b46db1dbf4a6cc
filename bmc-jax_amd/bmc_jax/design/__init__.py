"""The Cramer-Rao bound and protocol optimisation (src/bmc_jax/design/__init__.py)."""
from .crlb import crlb, CRLB, fisher_information, model_of
from .optimize import optimize_design, DesignResult

__all__ = ["crlb", "CRLB", "fisher_information", "model_of", "optimize_design", "DesignResult"]
