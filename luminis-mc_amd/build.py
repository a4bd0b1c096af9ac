"""Build libluminis_hip.so (gfx950 kernels + host table builders) in-tree.

Usage: python luminis-mc_amd/build.py   (also called by __graft_entry__.build)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "luminis_mc", "_lib")
LIB = os.path.join(OUT_DIR, "libluminis_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LMC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off: plain VALU FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose_isa: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    obj_t = os.path.join(build_dir, "lmc_transport.o")
    obj_h = os.path.join(build_dir, "lmc_host_tables.o")
    run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "lmc_transport.hip"), "-o", obj_t])
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, "lmc_host_tables.cpp"), "-o", obj_h])
    tmp = LIB + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", obj_t, obj_h, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    print(LIB)
