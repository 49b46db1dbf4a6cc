"""ctypes binding of libbmc_probe.so (csrc/probe/bmc_mfma_probe.hip): the squarings' dense product timed
two ways, the library's team column product and an FP64 MFMA tile product. A measurement tool beside the
simulate-and-fit ABI, not part of it."""
import ctypes
import os

from . import _native

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libbmc_probe.so")
PATH_COLS, PATH_MFMA, PATH_STAGED = 0, 1, 2
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise _native.NativeError(f"libbmc_probe.so is not built ({LIB_PATH}); run python bmc-jax_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.bmc_probe_products.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                         ctypes.c_int, ctypes.c_void_p]
        L.bmc_probe_products.restype = ctypes.c_int
        _lib = L
    return _lib


def products(path, nc, P, E, reps):
    """E <- P^reps E in place for every matrix of E [n, NA, NA] (torch float64 on the GPU), P [NA, NA]."""
    torch = _native.torch_cuda()
    na = 3 * (1 + nc) + 2
    if P.shape != (na, na) or E.dim() != 3 or E.shape[1:] != (na, na):
        raise ValueError(f"P must be {na} x {na} and E [n, {na}, {na}] for {nc} CEST pools")
    if P.dtype != torch.float64 or E.dtype != torch.float64 or not (P.is_cuda and E.is_cuda):
        raise ValueError("P and E must be float64 device tensors")
    if not (P.is_contiguous() and E.is_contiguous()):
        raise ValueError("P and E must be contiguous")
    rc = lib().bmc_probe_products(path, nc, P.data_ptr(), E.data_ptr(), E.shape[0], reps,
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise _native.NativeError(f"bmc_probe_products(path {path}, nc {nc}) failed with code {rc}")
    return E
