"""Build libcdet_hip.so (gfx950) in-tree with hipcc.

Usage: python cdet-renyi-entanglement-hpc_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "cdet_kernels.hip")
OUT = os.path.join(HERE, "libcdet_hip.so")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Reference parity: no FMA contraction (explicit fma only where the
    # reference calls std::fma).
    "-ffp-contract=off",
    # Keep packed FP32 instructions out of the code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: the default AGPR form copied every loop-
    # carried accumulator through v_accvgpr moves each K group.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
]


DRIVER_SRC = os.path.join(HERE, "driver", "cdet_mcmcmc.cpp")
DRIVER_OUT = os.path.join(HERE, "libcdet_mcmcmc.so")
# Host driver: g++/libstdc++ (the reference's toolchain), generic x86-64.
DRIVER_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-Wall"]


def build_driver(verbose=False):
    cxx = os.environ.get("CXX_HOST", "g++")
    tmp = DRIVER_OUT + ".tmp"
    cmd = [cxx] + DRIVER_FLAGS + ["-o", tmp, DRIVER_SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed")
    os.replace(tmp, DRIVER_OUT)
    return DRIVER_OUT


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, OUT)
    build_driver(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
