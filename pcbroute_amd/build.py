"""Builds libpcbgpu.so in-tree with hipcc for gfx950 (python -m pcbroute_amd.build)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpcbgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PCBGPU_ARCH", "gfx950")

# packed-fp32 instructions are switched off for every HIP translation unit
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-fPIC", "-Wall"]
SOURCES = ["pcbgpu.hip", "prior_cnn.hip", "backtrace.cpp", "gpu_shim.cpp", "router.cpp", "dsn.cpp"]
CHECK = os.path.join(HERE, "shim_check")  # a C++ program that uses the gpu.h shim (built beside the library)


def build_lib(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libpcbgpu.so")
    os.replace(tmp, LIB)
    cmd = ["g++", "-O2", "-std=c++17", "-o", CHECK, os.path.join(CSRC, "shim_check.cpp"), f"-L{HERE}", "-lpcbgpu",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for shim_check")
    return LIB


if __name__ == "__main__":
    print(build_lib(verbose=True))
