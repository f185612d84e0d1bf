"""Build libmfsv_hip.so (the HIP/gfx950 C-ABI library) in-tree.

    python -m mfcm_amd.build            # compile for gfx950
    python -m mfcm_amd.build --save-asm # also write the AMD assembly to _build/

Flags: -ffp-contract=off keeps the FP64 path's rounding identical to the
reference's operation order (fused operations only where the source writes
fma()); packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "mfsv.hip")
OUT = os.path.join(HERE, "libmfsv_hip.so")
DRIVER_SRC = os.path.join(HERE, "host", "mfsv_drive.cpp")
DRIVER = os.path.join(HERE, "host", "mfsv_drive")
PROBE_SRC = os.path.join(HERE, "csrc", "probe_march.hip")
PROBE = os.path.join(HERE, "probe_march")
ARCH = os.environ.get("MFSV_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose: bool = False, save_asm: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmfsv_hip.so")
    os.replace(tmp, OUT)
    if save_asm:
        bdir = os.path.join(os.path.dirname(HERE), "_build")
        os.makedirs(bdir, exist_ok=True)
        asm = os.path.join(bdir, "mfsv-gfx950.s")
        subprocess.run([hipcc, "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "--cuda-device-only", "-S", "-o", asm, SRC], check=True)
    build_driver(verbose)
    build_probe(verbose)
    return OUT


def build_probe(verbose: bool = False) -> str:
    """The memory-pattern probe (a measurement aid beside the library: the
    stage kernels' access pattern without their arithmetic)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-o", PROBE + ".tmp", PROBE_SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the memory-pattern probe failed")
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


def build_driver(verbose: bool = False) -> str:
    """The C++ mirror's command-line driver, linked against libmfsv_hip.so
    (found next to it through the run path)."""
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-o", DRIVER + ".tmp", DRIVER_SRC,
           f"-L{HERE}", "-lmfsv_hip", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building the C++ driver failed")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True, save_asm="--save-asm" in sys.argv)
    print("built", OUT)
