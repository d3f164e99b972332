"""Helpers for the tests of the command-line drivers (multigridmc_amd/bin)."""
import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "multigridmc_amd", "bin")
FIX = os.path.join(ROOT, "tests", "golden", "drivers")


def run_driver(name, tmp_path, cfg_text=None, meas_text=None, timeout=600):
    """Copy the fixtures into tmp_path (optionally replacing their text), run
    the driver there and return (returncode, stdout)."""
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build with python -m multigridmc_amd.build")
    for f in ("posterior_2d.cfg", "measurements_2d.cfg"):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    if cfg_text is not None:
        (tmp_path / "posterior_2d.cfg").write_text(cfg_text)
    if meas_text is not None:
        (tmp_path / "measurements_2d.cfg").write_text(meas_text)
    p = subprocess.run([exe, "posterior_2d.cfg"], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def fixture_text(name):
    with open(os.path.join(FIX, name)) as fh:
        return fh.read()
