"""The C ABI library loads and exports every entry point include/bmc_hip.h declares (CPU, no compute)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bmc_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bmc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("bmc_abi_version", "bmc_last_error", "bmc_plan_create", "bmc_simulate", "bmc_lm_fit"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from bmc_jax import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libbmc_hip.so is not built (python bmc-jax_amd/build.py)")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    assert lib.bmc_abi_version() == 3


def test_errors_are_reported_without_a_device():
    """Argument errors come back as codes with a message (no GPU needed)."""
    from bmc_jax import _native
    L = _native.lib()
    h = ctypes.c_void_p()
    rc = L.bmc_plan_create_ex(0, 1, None, None, None, None, None, 1, 1, ctypes.byref(h))
    assert rc == -1 and b"bad arguments" in L.bmc_last_error()
    assert L.bmc_plan_destroy(None) == 0


def test_expm_argument_errors_without_a_device():
    """bmc_expm refuses sizes outside 1..BMC_EXPM_MAX_N and unknown dtypes before touching the device."""
    from bmc_jax import _native
    L = _native.lib()
    assert L.bmc_expm(None, None, 25, 1, 1, None) == -1 and b"n must be" in L.bmc_last_error()
    assert L.bmc_expm(None, None, 0, 1, 1, None) == -1
    assert L.bmc_expm(None, None, 4, 1, 7, None) == -1 and b"dtype" in L.bmc_last_error()
    assert L.bmc_expm(None, None, 4, 0, 1, None) == 0
