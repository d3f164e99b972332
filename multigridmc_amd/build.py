"""Builds libmgmc_amd.so in-tree for gfx950 (hipcc) from multigridmc_amd/csrc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmgmc_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 (v_pk_*_f32) code generation is disabled for every HIP file.
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def build(verbose=True):
    obj = os.path.join(CSRC, "setup.o")
    oobj = os.path.join(CSRC, "operators.o")
    cobj = os.path.join(CSRC, "cholesky_host.o")
    hobj = os.path.join(CSRC, "mgmc_lib.o")
    cmds = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, "setup.cc"), "-o", obj],
        ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, "operators.cc"), "-o", oobj],
        # host Cholesky samplers: AVX2/FMA (every x86-64 server CPU since 2013)
        ["g++", "-O3", "-mavx2", "-mfma", "-std=c++17", "-fPIC", "-Wall", "-Wno-unknown-pragmas", "-c", os.path.join(CSRC, "cholesky_host.cc"),
         "-o", cobj],
        [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "mgmc_lib.hip"), "-o", hobj],
        [HIPCC, "--offload-arch=gfx950", "-shared", hobj, obj, oobj, cobj, "-lpthread", "-o", LIB + ".tmp"],
    ]
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        subprocess.check_call(c)
    os.replace(LIB + ".tmp", LIB)
    build_drivers(verbose)
    return LIB


ADAPTERS = os.path.join(HERE, "libmgmc_adapters.so")


def build_adapters(verbose=True):
    """libmgmc_adapters.so: the C++ classes with the reference's Sampler /
    Preconditioner / LoopSolver shapes over the C ABI, and bin/adapter_demo."""
    src = os.path.join(CSRC, "adapters")
    c = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", os.path.join(src, "mgmc_adapters.cc"),
         "-L" + HERE, "-lmgmc_amd", "-Wl,-rpath,$ORIGIN", "-o", ADAPTERS + ".tmp"]
    if verbose:
        print(" ".join(c), flush=True)
    subprocess.check_call(c)
    os.replace(ADAPTERS + ".tmp", ADAPTERS)
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "adapter_demo")
    c = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(src, "adapter_demo.cc"), "-L" + HERE,
         "-lmgmc_adapters", "-lmgmc_amd", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(c), flush=True)
    subprocess.check_call(c)
    os.replace(out + ".tmp", out)


def build_drivers(verbose=True):
    """driver_mgmc and driver_mg (host C++ over the C ABI; driver_mg through
    the adapter classes) into multigridmc_amd/bin."""
    build_adapters(verbose)
    host = os.path.join(CSRC, "host")
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    common = [os.path.join(host, "config.cc"), os.path.join(host, "driver_common.cc")]
    for name in ("driver_mgmc", "driver_mg"):
        out = os.path.join(bindir, name)
        c = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(host, name + ".cc")] + common + [
            "-L" + HERE, "-lmgmc_adapters", "-lmgmc_amd", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(c), flush=True)
        subprocess.check_call(c)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
