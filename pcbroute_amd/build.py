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
SOURCES = ["pcbgpu.hip", "prior_cnn.hip", "backtrace.cpp", "chain.cpp", "astar.cpp", "gpu_shim.cpp", "router.cpp",
           "dsn.cpp"]
OBJ = os.path.join(HERE, "_build")       # object files (not shipped in the history)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode == 0


def build_lib(verbose=False):
    """Every source compiled on its own (in parallel), then linked: a change to one file recompiles all of them
    (no dependency tracking), but the slow HIP files no longer wait for each other."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in srcs]
    jobs = [[HIPCC, *HIP_FLAGS, "-c", "-o", o, s] for s, o in zip(srcs, objs)]
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        ok = list(ex.map(lambda c: _run(c, verbose), jobs))
    if not all(ok):
        raise RuntimeError("hipcc failed for libpcbgpu.so")
    tmp = LIB + ".tmp"
    if not _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs, "-lpthread"], verbose):
        raise RuntimeError("hipcc failed to link libpcbgpu.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_lib(verbose=True))
