"""MR-fingerprinting dictionaries from the simulator (src/bmc_jax/mrf/__init__.py)."""
from .dictionary import build_param_grid, generate_dictionary

__all__ = ["build_param_grid", "generate_dictionary"]
