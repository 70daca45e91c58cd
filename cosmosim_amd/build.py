"""Build libcosmosim.so in-tree: the HIP core (gfx950) and the C host layer.

    python -m cosmosim_amd.build          # or cosmosim_amd.build.build()
    python -m cosmosim_amd.build --ic-source /path/to/src/initial_conditions.c [--out LIB]

hipcc compiles csrc/csg_core.hip for gfx950 only; gcc compiles the C host
layer; hipcc links both into cosmosim_amd/libcosmosim.so (same name as the
reference's shared library, so consumers that dlopen it find it unchanged).

--ic-source links the caller's own initial-condition generators into the
library, compiled where the file lies with its own directory first on the
include path, as the reference's shared-library target links
src/initial_conditions.c (CMakeLists.txt:55-58).  cosmosim_create needs them:
without them it returns NULL (COSMOSIM_E_UNSUPPORTED) rather than generate
other initial conditions.  The default build also links the test-only
generators of tests/ic_stub.c into tests/_lib/libcosmosim_ic_stub.so for the
handle tests.
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


IC_STUB = os.path.join(ROOT, "tests", "ic_stub.c")
IC_STUB_LIB = os.path.join(ROOT, "tests", "_lib", "libcosmosim_ic_stub.so")


def _objects():
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
    return objs


def _link(objs, out, ic_sources=(), ic_includes=()):
    """Link the core objects (and the given generator sources) into `out`."""
    extra = []
    for k, src in enumerate(ic_sources):
        src = os.path.abspath(src)
        o = os.path.join(BUILD, f"ic{k}_" + os.path.basename(src) + ".o")
        incs = ["-I" + os.path.dirname(src)] + ["-I" + i for i in ic_includes]
        # the caller's file as its own build compiles it (gnu11: M_PI)
        _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", *incs, "-c", src,
              "-o", o])
        extra.append(o)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp,
          *objs, *extra, "-lm"])
    shutil.move(tmp, out)
    return out


def build(verbose=False, ic_sources=(), out=LIB, stub=True):
    """Compile the library; ic_sources: the caller's initial-condition generators
    (C files defining generate_spiral_galaxy, generate_merger, ...).  With stub,
    the test variant with tests/ic_stub.c is linked from the same objects."""
    objs = _objects()
    lib = _link(objs, out, ic_sources)
    if verbose:
        print("built", lib + (" with " + ", ".join(ic_sources) if ic_sources else ""))
    if stub and os.path.exists(IC_STUB):
        _link(objs, IC_STUB_LIB, [IC_STUB], [INC])
        if verbose:
            print("built", IC_STUB_LIB)
    return lib




if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="build libcosmosim.so for gfx950")
    ap.add_argument("--ic-source", action="append", default=[],
                    help="C file with the caller's initial-condition generators")
    ap.add_argument("--out", default=LIB)
    a = ap.parse_args()
    build(verbose=True, ic_sources=a.ic_source, out=a.out, stub=a.out == LIB)
