"""Build libcosmosim.so in-tree: the HIP core (gfx950) and the C host layer.

    python -m cosmosim_amd.build          # or cosmosim_amd.build.build()

hipcc compiles csrc/csg_core.hip for gfx950 only; gcc compiles the C host
layer; hipcc links both into cosmosim_amd/libcosmosim.so (same name as the
reference's shared library, so consumers that dlopen it find it unchanged).
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcosmosim.so")
ARCH = os.environ.get("COSMOSIM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["csg_core.hip"]
C_SOURCES = ["legacy.c", "api.c", "ic_seeded.c"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + err)
    if err.strip():
        print(err, file=sys.stderr)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I" + INC, "-I" + CSRC, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in C_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run(["gcc", *C_FLAGS, "-I" + INC, "-I" + CSRC, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lm"])
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
