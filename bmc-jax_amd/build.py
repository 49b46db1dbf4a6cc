"""Builds libbmc_hip.so (the HIP kernels and the C ABI) in-tree, for gfx950.

    python bmc-jax_amd/build.py            # compile every .hip source and link the shared library
    python bmc-jax_amd/build.py --incremental   # recompile only sources older than their inputs

The library lands in bmc-jax_amd/bmc_jax/_lib/, where the Python layer loads it.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "bmc_jax", "_lib")
LIB = os.path.join(OUT_DIR, "libbmc_hip.so")
PROBE_LIB = os.path.join(OUT_DIR, "libbmc_probe.so")
SOURCES = ["bmc_abi.hip", "bmc_sim.hip", "bmc_cwj.hip", "bmc_adj.hip", "bmc_uneven.hip", "bmc_lm.hip",
           "bmc_mlp.hip", "bmc_expm.hip"]
ARCH = os.environ.get("BMC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _stale(obj, src):
    """True when obj is missing or older than its source or any header it may include."""
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(ROOT, "include", "bmc_hip.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, jobs: int = 4, defines=(), incremental: bool = False) -> str:
    """Compile every source (or, with incremental, only those older than their inputs) and link."""
    os.makedirs(OUT_DIR, exist_ok=True)
    objs_dir = os.path.join(HERE, "_build")
    os.makedirs(objs_dir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objs_dir, src.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and not defines and not _stale(obj, os.path.join(CSRC, src)):
            continue
        cmd = [hipcc()] + HIPCC_FLAGS + [f"-D{d}" for d in defines] + inc + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        noise = "is not a recognized feature for this target"
        lines = [l for l in text.splitlines() if noise not in l]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_probe(incremental=incremental)
    return LIB


def build_probe(incremental: bool = False) -> str:
    """libbmc_probe.so: the product-path measurement probe (csrc/probe), apart from the ABI library."""
    src = os.path.join(CSRC, "probe", "bmc_mfma_probe.hip")
    if incremental and not _stale(PROBE_LIB, src):
        return PROBE_LIB
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    tmp = PROBE_LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + inc + ["-shared", src, "-o", tmp]
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if out.returncode != 0:
        print(out.stdout.decode(errors="replace"), file=sys.stderr)
        raise RuntimeError("hipcc failed for the probe library")
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
