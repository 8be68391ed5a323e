"""Builds libdkt.so in-tree (host mesh layer with g++, HIP kernels with hipcc
for gfx950).  Run as ``python dendro-kt_amd/build.py`` or through
__graft_entry__.build().  Always recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdkt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DKT_OFFLOAD_ARCH", "gfx950")

HOST_SRCS = ["mesh.cpp", "capi_mesh.cpp"]
HIP_SRCS = ["plan.hip", "cg.hip", "batch.hip"]

HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-variable",
              "-Wno-sign-compare"]
# Packed FP32 instructions are disabled for all device code (some of their forms have
# given wrong results on MI355X without any error).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=fast",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build():
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(OUT, s + ".o")
        run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(OUT, s + ".o")
        run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-L/usr/lib/gcc/x86_64-linux-gnu/11",
         "-lgomp", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
