"""Build libluminis_hip.so (gfx950 kernels + host table builders) in-tree.

Usage: python luminis-mc_amd/build.py   (also called by __graft_entry__.build)
"""
import os
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


N_PARTS = 5  # lmc_transport_kernels.hip parts (LMC_TK_PART)


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def run_parallel(cmds, jobs):
    """Run compile commands, at most `jobs` at a time; raise on the first failure."""
    pending, running = list(cmds), []
    while pending or running:
        while pending and len(running) < jobs:
            cmd = pending.pop(0)
            print("+", " ".join(cmd), flush=True)
            running.append((subprocess.Popen(cmd), cmd))
        proc, cmd = running.pop(0)
        if proc.wait() != 0:
            for p, _ in running:
                p.wait()
            raise subprocess.CalledProcessError(proc.returncode, cmd)


def build(csrc: str = CSRC, out_dir: str = OUT_DIR, build_dir: str | None = None) -> str:
    """Compile the host TU, the transport-kernel parts and the host table
    builders (in parallel), then link libluminis_hip.so into out_dir."""
    os.makedirs(out_dir, exist_ok=True)
    build_dir = build_dir or os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs, cmds = [], []
    obj = os.path.join(build_dir, "lmc_transport.o")
    cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(csrc, "lmc_transport.hip"), "-o", obj])
    objs.append(obj)
    for part in range(N_PARTS):
        obj = os.path.join(build_dir, f"lmc_transport_kernels_{part}.o")
        cmds.append([HIPCC, *HIP_FLAGS, f"-DLMC_TK_PART={part}", "-c",
                     os.path.join(csrc, "lmc_transport_kernels.hip"), "-o", obj])
        objs.append(obj)
    obj = os.path.join(build_dir, "lmc_host_tables.o")
    cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(csrc, "lmc_host_tables.cpp"), "-o", obj])
    objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    run_parallel(cmds, jobs)
    lib = os.path.join(out_dir, "libluminis_hip.so")
    tmp = lib + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build()
    print(LIB)
