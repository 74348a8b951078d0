"""Builds gpview_amd/libgpv.so (HIP for gfx950 + host runtime) in-tree.

    python -m gpview_amd.build          # or: python gpview_amd/build.py

Flags: -ffp-contract=off and no fast-math on host and device (occupancy is an
exact quantity; no multiply/add pair may be fused), packed-fp32 instructions
switched off, correctly rounded f32 division (the HIP default, checked by
tests/test_abi.py::test_codegen_exactness on the disassembly).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgpv.so")

SOURCES = [
    "gpv_kernels.hip",
    "gpv_runtime.cpp",
    "gpv_mesh.cpp",
    "gpv_object.cpp",
    "gpv_serve.cpp",
    "gpv_probe.hip",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", INCLUDE, "-I", CSRC,
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgpv")


def build(verbose=False, out=None, defines=(), objdir=None):
    """Compiles the library to `out` (default gpview_amd/libgpv.so); defines
    are extra -D macros for alternative builds of the same sources."""
    cc = hipcc()
    objdir = objdir or os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    extra = ["-D" + d for d in defines]
    cmds = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        cmds.append([cc, "-x", "hip", "-c", path, "-o", obj, "-pthread"] + COMMON + extra)
        objs.append(obj)
    from concurrent.futures import ThreadPoolExecutor

    def run(cmd):
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1, 8)) as ex:
        list(ex.map(run, cmds))
    lib = out or LIB
    tmp = lib + ".tmp"
    cmd = [cc, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-fPIC", "-pthread"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


def disassemble():
    """AMD ISA of our own kernels (for the codegen checks)."""
    objdir = os.path.join(HERE, "_build")
    path = os.path.join(CSRC, "gpv_kernels.hip")
    out = os.path.join(objdir, "gpv_kernels.s")
    cmd = [hipcc(), "-x", "hip", "-S", "--cuda-device-only", path, "-o", out] + COMMON
    subprocess.run(cmd, check=True)
    with open(out) as f:
        return f.read()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
