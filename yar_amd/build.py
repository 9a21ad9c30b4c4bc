"""Build recipe for libyar_gpu.so: host C++ (g++) + HIP kernels for gfx950 (hipcc).

Run as ``python -m yar_amd.build``; objects go to yar_amd/_build/, the library to
yar_amd/libyar_gpu.so (in-tree, so it travels with the source tree).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libyar_gpu.so")

HOST_SOURCES = ["host/common.cpp", "host/pbrt_loader.cpp", "host/kdtree_builder.cpp", "host/scene_gen.cpp",
                "host/checkpoint.cpp", "host/exr_writer.cpp"]
HIP_SOURCES = ["gpu/render.hip", "gpu/context.hip", "gpu/guide.hip"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
CXX = shutil.which("g++") or "g++"

# -ffp-contract=off everywhere: the path keeps the reference's unfused float arithmetic.
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in HOST_SOURCES + HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs + ["-pthread", "-lz"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
