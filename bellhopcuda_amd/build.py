"""Builds libbhc_hip.so (HIP kernels for gfx950 + host C++ front end) in-tree.

Flags: -O3, -ffp-contract=off on every translation unit (trajectories are
compared bit for bit with the reference's CPU arithmetic, which never fuses
a*b+c), and packed-FP32 instructions disabled for device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbhc_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BHC_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["trace2d.hip", "mlp.hip", "opcount2d.cpp"]
CXX_SOURCES = ["host2d.cpp", "host3d.cpp", "writers2d.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-value"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run(["g++"] + COMMON + ["-Wall", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    # Link with g++ so that complex division / multiplication in the host code
    # resolve to libgcc's __divdc3 / __muldc3, as in the reference's g++ build
    # (clang's link would bind compiler-rt's versions, which round differently).
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    _run(["g++", "-shared", "-fPIC", "-o", tmp] + objs
         + ["-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(rocm, "lib")])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
