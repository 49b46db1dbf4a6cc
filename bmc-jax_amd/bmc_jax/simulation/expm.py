"""One time step of the Bloch-McConnell system, for callers that hold A and C themselves.

The reference's time step (src/bmc_jax/simulation/expm.py:86-119) is M <- [exp([[A, C], [0, 0]] dt) [M; 1]]_N.
Inside a simulation the HIP kernel takes this step itself (csrc/bmc_sim.hip). These functions serve callers
that build A and C by hand; the exponential is the library's bmc_expm kernel (csrc/bmc_expm.hip): Pade
approximation with scaling and squaring, float64 orders 3 to 13 for float64 input, float32 orders 3/5/7
computed in float32 for float32 input (the reference's expm_library and expm_plain, expm.py:26-83).
"""
import numpy as np

from .. import _native, settings


def expm_library(A):
    """exp(A) on the GPU (expm.py:26-28), in A's precision; A may carry leading batch axes."""
    return _native.expm(np.asarray(A))


def expm_plain(A, max_squarings=16):
    """exp(A) by the float32 algorithm (expm.py:50-83): Pade 3/5/7 by the unscaled 1-norm, Gauss-Jordan
    solve, NaN past 16 squarings. float32 only, as the reference's."""
    a = np.asarray(A)
    if a.dtype != np.float32:
        raise ValueError('use_expm("plain") takes float32 matrices only; x64 is on or the input is float64')
    if int(max_squarings) != 16:
        raise ValueError("the kernel's squaring limit is 16, the reference's default")
    return _native.expm(a)


def propagate(M, dt, A, C):
    """Solve dM/dt = A M + C over dt (expm.py:115-119): one augmented exponential and a matvec."""
    M, A, C = np.asarray(M), np.asarray(A), np.asarray(C)
    n = M.shape[0]
    aug = np.zeros((n + 1, n + 1), dtype=A.dtype)
    aug[:n, :n] = A
    aug[:n, n] = C
    E = (expm_plain if settings.expm_backend() == "plain" else expm_library)(aug * np.asarray(dt, A.dtype))
    return (E @ np.append(M, 1.0).astype(E.dtype))[:n]
