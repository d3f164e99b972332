"""The library's host set-up code (hierarchy construction, Galerkin
stencils, SMW factors, lexicographic capacitance matrices, measurement
vectors, the CSR-prior path) built with AddressSanitizer and
UndefinedBehaviorSanitizer by g++ and run over 2D/3D cases, including x
extents that are multiples of 8 (padded rows share their boundary slots).
CPU only."""
import os
import subprocess

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
