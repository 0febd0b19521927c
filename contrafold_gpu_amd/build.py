"""Build the gfx950 fold engine: one shared library, libcontrafold_amd.so.

    python -m contrafold_gpu_amd.build

hipcc compiles the HIP sources for gfx950 only, with -ffp-contract=off (every
multiply-add is two rounded operations, as in the reference's float32 engine)
with packed-FP32 instructions disabled, and with -fno-honor-nans (no value
the kernels handle is ever NaN -- an empty cell is -1e30 -- so min/max need no
NaN quieting).  The library is written in-tree
next to this file so that it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libcontrafold_amd.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["cf_api.hip", "cf_server.hip"]
DEPS = ["cf_api.hip", "cf_server.hip", "cf_kernels.hip", "cf_device.h", "cf_internal.h"]
ARCH = os.environ.get("CF_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the fold engine needs the ROCm toolchain to build")


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-fno-honor-nans",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(os.path.dirname(HERE), "include"), "-I" + CSRC,
            "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


def build(force=True, verbose=False):
    """Compile every HIP source and link the shared library.  Always rebuilds
    when force is true (the build check compiles every time it is asked)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    if not force and os.path.exists(LIB_PATH):
        newest = max(os.path.getmtime(os.path.join(CSRC, d)) for d in DEPS if os.path.exists(os.path.join(CSRC, d)))
        if os.path.getmtime(LIB_PATH) >= newest:
            return LIB_PATH
    tmp = LIB_PATH + ".tmp"
    objs = []
    for src in srcs:
        obj = os.path.join(CSRC, os.path.basename(src) + ".o")
        cmd = [hipcc()] + flags() + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    return LIB_PATH


if __name__ == "__main__":
    print(build(force=True, verbose="-v" in sys.argv))
