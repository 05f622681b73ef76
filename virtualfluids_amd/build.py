"""Build the gfx950 shared library libvfamd.so in-tree with hipcc.

Every translation unit under csrc/ (HIP kernels, C entry points and the C++ host
classes) is compiled for gfx950 and linked into ``virtualfluids_amd/libvfamd.so``.
Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvfamd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ROOT = os.path.dirname(HERE)
HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_host.cpp")
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "test_host")
K17_HOST_SRC = os.path.join(ROOT, "tests", "cpp", "k17_host.cpp")
K17_HOST_LIB = os.path.join(ROOT, "tests", "cpp", "libk17host.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "host", "*.cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.relpath(src, CSRC).replace(os.sep, "_") + ".o")
    cmd = [HIPCC, *DEVICE_FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *DEVICE_FLAGS, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile every source (always, so a stale library is never reused) and link."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host_test()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources (and {HOST_TEST})", file=sys.stderr)
    return LIB


def build_host_test() -> str:
    """The C++ host-layer test program, linked against the in-tree libvfamd.so."""
    tmp = HOST_TEST + ".tmp"
    cmd = [HIPCC, *DEVICE_FLAGS, "-x", "hip", HOST_TEST_SRC, "-o", tmp, f"-L{HERE}", "-lvfamd",
           "-Wl,-rpath,$ORIGIN/../../virtualfluids_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {HOST_TEST_SRC}:\n{r.stderr}")
    os.replace(tmp, HOST_TEST)
    # the collision and its adjoint compiled for the host (CPU tests of the device source)
    tmp = K17_HOST_LIB + ".tmp"
    cmd = [HIPCC, *DEVICE_FLAGS, "-x", "hip", "-shared", K17_HOST_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {K17_HOST_SRC}:\n{r.stderr}")
    os.replace(tmp, K17_HOST_LIB)
    return HOST_TEST


if __name__ == "__main__":
    build()
