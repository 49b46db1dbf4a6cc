"""One time step of the Bloch-McConnell system, for callers that hold A and C themselves.

The reference's time step (src/bmc_jax/simulation/expm.py:86-119) is M <- [exp([[A, C], [0, 0]] dt) [M; 1]]_N.
Inside a simulation the HIP kernel takes this step itself (csrc/bmc_sim.hip). These functions serve
callers that build A and C by hand: the exponential runs on the GPU through torch.linalg.matrix_exp
(a library routine; `use_expm` only keeps the reference's float32-only rule for "plain").
"""
import numpy as np

from .. import _native, settings


def expm_library(A):
    """Matrix exponential on the GPU (expm.py:26-28); A may carry leading batch axes."""
    torch = _native.torch_cuda()
    a = np.asarray(A)
    t = torch.as_tensor(np.ascontiguousarray(a)).to("cuda")
    return torch.linalg.matrix_exp(t).cpu().numpy()


def expm_plain(A, max_squarings=16):
    """As expm_library; float32 only, like the reference's "plain" exponential (expm.py:50-54)."""
    if np.asarray(A).dtype != np.float32:
        raise ValueError('use_expm("plain") is float32 only (x64 is on)')
    return expm_library(A)


def propagate(M, dt, A, C):
    """Solve dM/dt = A M + C over dt (expm.py:115-119): one augmented exponential and a matvec."""
    M, A, C = np.asarray(M), np.asarray(A), np.asarray(C)
    n = M.shape[0]
    aug = np.zeros((n + 1, n + 1), dtype=A.dtype)
    aug[:n, :n] = A
    aug[:n, n] = C
    E = (expm_plain if settings.expm_backend() == "plain" else expm_library)(aug * np.asarray(dt, A.dtype))
    return (E @ np.append(M, 1.0).astype(E.dtype))[:n]
