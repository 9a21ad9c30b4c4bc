"""Builds the CPU oracle (test infrastructure, never part of the product library).

  oracle/liboracle.so      - cpu_ref.cpp, the scalar restatement of the RAY hot path
  oracle/_ref/libpcg_ref.so - the reference's own third-party/pcg/pcg_basic.c, compiled unchanged
                             from /root/reference when that tree is present (it is the only part of
                             the reference path that builds without stand-ins for MSVC intrinsics)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE = "/root/reference"
PCG_SRC = os.path.join(REFERENCE, "third-party", "pcg", "pcg_basic.c")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + r.stdout)


def build():
    out = os.path.join(HERE, "liboracle.so")
    # -O2 -mavx2 -mfma -ffp-contract=off: the reference probe's settings, without FMA contraction
    _run(["g++", "-std=c++17", "-O2", "-mavx2", "-mfma", "-ffp-contract=off", "-fPIC", "-shared", "-pthread",
          os.path.join(HERE, "cpu_ref.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    built = [out]
    if os.path.exists(PCG_SRC):
        ref = os.path.join(HERE, "_ref", "libpcg_ref.so")
        os.makedirs(os.path.dirname(ref), exist_ok=True)
        _run(["gcc", "-O2", "-fPIC", "-shared", PCG_SRC, "-o", ref + ".tmp"])
        os.replace(ref + ".tmp", ref)
        built.append(ref)
    return built


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
