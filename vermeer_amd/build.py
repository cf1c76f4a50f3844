"""Builds the HIP library (libvermeer_gpu.so) in-tree for gfx950.

Every source is compiled with -ffp-contract=off: the traversal must round
each float operation exactly as the reference's unfused SSE code, and the
builder's box arithmetic likewise.  Packed-FP32 instructions are disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvermeer_gpu.so")

SOURCES = ["trace.hip", "raysort.hip", "hbmcopy.hip", "output.hip", "vnf.cpp", "render.hip", "guide.hip", "pathtrace.hip", "scene.cpp", "render.cpp", "guide.cpp",
           "pathtrace.cpp", "qbvh_build.cpp", "bvh_build.hip"]
HEADERS = ["trace_kernels.h", "render_kernels.h", "guide_kernels.h", "guide_internal.h", "pathtrace_kernels.h",
           "scene_internal.h", "qbvh_build.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    """Compiles all sources and links LIB. Always rebuilds."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        cmds.append([HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", path, "-o", obj])
        objs.append(obj)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
