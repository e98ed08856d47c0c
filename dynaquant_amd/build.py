"""Build libdynaquant_amd.so (gfx950) in-tree with hipcc.

Usage: python -m dynaquant_amd.build  (or dynaquant_amd.build.build()).
Compiles every HIP source on each call; the CPU-only container
cross-compiles for gfx950 without a GPU.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdynaquant_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# no packed-FP32 instructions: the engine's FP32 arithmetic stays scalar
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]

SOURCES_HIP = ["dq_kernels.hip", "dq_batched.hip", "dq_ring.hip", "dq_adjoint.hip", "dq_noalloc.hip"]
SOURCES_CPP = ["dq_api.cpp", "dq_tube.cpp", "dq_io.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES_HIP:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, "-x", "hip", "--offload-arch=" + ARCH] + COMMON + DEVICE_FLAGS +
             ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in SOURCES_CPP:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
