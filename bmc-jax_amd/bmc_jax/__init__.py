"""bmc_jax on MI355X: the BMC-JAX simulate-and-fit path on hand-written CDNA4 HIP kernels."""
__version__ = "0.2.0"
