"""Build the walker engine library (``libmqe.so``) for gfx950 with hipcc.

The library is built in-tree so that it travels with the repository; the
package loads it from ``mpsqmc_amd/libmqe.so``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmqe.so")

SOURCES = ["mqe_engine.hip", "mqe_mlp.hip", "mqe_noalloc.hip", "mqe_host.cpp", "mqe_dmrg.cpp", "mpsqmc.cpp"]
WALK = os.path.join(HERE, "mpsqmc_walk")
TOOL = os.path.join(HERE, "mpsqmc_tool")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    # one object per source, compiled concurrently (the walker engine's device
    # code takes the longest), then one link
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"] + ["-c"]
    jobs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + compile_flags + inc + [os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [cmd for cmd, p in jobs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    # the command-line walk driver over the C++ host classes
    cmd2 = [hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "mpsqmc_main.cpp"),
            "-L", HERE, "-lmqe", "-Wl,-rpath,$ORIGIN", "-o", WALK + ".tmp"]
    if verbose:
        print(" ".join(cmd2), flush=True)
    subprocess.check_call(cmd2)
    os.replace(WALK + ".tmp", WALK)
    # one MPSstate / Walker method per call, for the host-class tests
    cmd3 = [hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "mpsqmc_tool.cpp"),
            "-L", HERE, "-lmqe", "-Wl,-rpath,$ORIGIN", "-o", TOOL + ".tmp"]
    if verbose:
        print(" ".join(cmd3), flush=True)
    subprocess.check_call(cmd3)
    os.replace(TOOL + ".tmp", TOOL)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
    sys.exit(0)
