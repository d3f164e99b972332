"""Sanitizer and emulation tests on the CPU.

- The library's host set-up code (hierarchy construction, Galerkin stencils,
  SMW factors, lexicographic capacitance matrices, measurement vectors, the
  CSR-prior path) built with AddressSanitizer and UndefinedBehaviorSanitizer
  by g++ and run over 2D/3D cases, including x extents that are multiples of
  8 (padded rows share their boundary slots).
- The kernels themselves, compiled with -DMGMC_EMULATION against the host
  runtime of tests/emu/hip_emu.h, run under the same sanitizers and checked
  against the oracle.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "multigridmc_amd", "csrc")


def test_setup_code_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "setup_check")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(HERE, "asan", "setup_check.cc"), os.path.join(CSRC, "setup.cc"),
           os.path.join(CSRC, "operators.cc"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "setup_check ok" in r.stdout


EMU = os.path.join(HERE, "emu")
LIB_SOURCES = [os.path.join(CSRC, "mgmc_lib.hip")]
HOST_SOURCES = [os.path.join(CSRC, f) for f in ("setup.cc", "operators.cc", "cholesky_host.cc")]


def test_kernels_under_emulation_asan_and_ubsan(tmp_path):
    """The library's kernels run on the CPU (tests/emu/hip_emu.h: one fiber
    per GPU thread, malloc'd device memory and LDS) under AddressSanitizer and
    UndefinedBehaviorSanitizer: sampling cycles and multigrid solves in FP64 and
    FP32, multicolour and lexicographic order, constant and variable stencils,
    x extents that are and are not multiples of 8, and a batch of seven 3D and 2D
    instances of different sizes and chain counts whose merged launches (plain and persistent)
    must reproduce every instance's own run bit for bit (tests/emu/emu_cycle.cc).
    Any out-of-bounds device or LDS access aborts the run."""
    exe = str(tmp_path / "emu_cycle")
    cmd = (["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
            "-DMGMC_EMULATION", "-I", EMU, "-x", "c++"] + LIB_SOURCES + ["-x", "none"] + HOST_SOURCES +
           [os.path.join(EMU, "emu_cycle.cc"), "-o", exe])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "emu_cycle ok" in r.stdout


def test_kernel_parity_with_the_oracle_in_emulation(tmp_path):
    """The same kernels (emulated on the CPU) against the oracle in FP64:
    operator, sweeps with and without injected noise on every level,
    residual + restriction, prolongation, coarsest-level solve and sample, to
    1e-12 (tests/emu/parity_check.py)."""
    so = str(tmp_path / "libmgmc_emu.so")
    cmd = (["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-DMGMC_EMULATION", "-I", EMU, "-x", "c++"] +
           LIB_SOURCES + ["-x", "none"] + HOST_SOURCES + ["-o", so])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([sys.executable, os.path.join(EMU, "parity_check.py"), so], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    worst = float([ln for ln in r.stdout.splitlines() if ln.startswith("worst")][0].split()[1])
    print(r.stdout)
    assert worst < 1e-12
