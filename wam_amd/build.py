"""Builds the in-tree HIP library libwamgpu.so for gfx950.

The HIP sources are compiled with hipcc; -ffp-contract=off keeps every
multiply and add separately rounded (the kernels reproduce the reference's
REAL*4 operation order), division and square root are correctly rounded,
and packed FP32 instructions are disabled.  Host code is built with -mfma so
that the double-precision fused multiply-adds of wam_math.h are single
instructions (the same correctly rounded operation as the library call).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwamgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WAM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["wam_prop.hip", "wam_implsch.hip", "wam_adjoint.hip", "wam_output.hip", "wam_nest.hip", "wam_capi.hip"]
CPP_SOURCES = ["wam_setup.cpp", "wam_restart.cpp", "wam_boundary_file.cpp", "wam_map_file.cpp", "wam_checkpoint.cpp"]
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Xarch_host", "-mfma", "-I" + CSRC]
CPP_FLAGS = ["-O2", "-mfma", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I" + CSRC]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in CPP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        _run(["g++"] + CPP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
