"""The C++ adapter classes with the reference's interface shapes
(multigridmc_amd/csrc/adapters: HipMultigridMCSampler : Sampler,
HipMultigridPreconditioner : Preconditioner, HipLoopSolver;
src/sampler/sampler.hh:23-78, src/preconditioner/preconditioner.hh:17-37,
src/solver/loop_solver.cc:9-54), driven by bin/adapter_demo: its results
equal the Python binding's bit for bit on the same handles' parameters,
the solve matches the oracle, fix_rhs leaves the sample path unchanged and
deep_copy with another generator gives another chain."""
import os
import subprocess

import numpy as np
import pytest

import multigridmc_amd as mg
from gpu_common import oracle_mg, params
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "multigridmc_amd", "bin")


def test_adapter_classes_through_adapter_demo(tmp_path):
    print("case: adapter_demo, 2D 32^2 posterior, 3 balls", flush=True)
    p = subprocess.run([os.path.join(BIN, "adapter_demo"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    out = dict(ln.split(" ", 1) for ln in p.stdout.splitlines() if ln.startswith("key"))
    assert "length error status 1" in p.stdout and "adapter_demo done" in p.stdout
    ld = lambda name: np.fromfile(tmp_path / name, dtype=np.float64)  # noqa: E731
    lat = mg.Lattice(32, 32)
    centres = np.array([[0.3, 0.4], [0.7, 0.55], [0.45, 0.75]])
    var = np.array([1e-3, 2e-3, 1.5e-3])
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.1, var)
    N = lat.Nvertex
    par = mg.MultigridParameters(nlevel=4)
    b = ld("b.bin")
    # Preconditioner::apply and LoopSolver::apply
    pre = mg.MultigridPreconditioner(op, par)
    assert np.array_equal(ld("precondition.bin"), pre.apply(b))
    solver = mg.LoopSolver(op, pre, rtol=1e-11, atol=1e-9, maxiter=50)
    x = solver.apply(b)
    assert np.array_equal(ld("solve.bin"), x)
    assert np.array_equal(ld("history.bin"), solver.history)
    ora = po.Operator.fd(2, (32, 32), 10.0)
    ora.add_measurements(centres, 0.1, var)
    x_ref, _, _ = oracle_mg(ora, params(4), ordering=po.LEX).loop_solve(b, rtol=1e-12, atol=1e-12, maxiter=80)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-10
    # Sampler::apply: the same chain as the Python binding with the key the
    # adapter drew from its generator
    f = ld("f.bin")
    y = np.array([0.5, -1.0, 0.25])
    assert np.abs(f - op.get_B() @ (y / op.get_Sigma())).max() <= 1e-12 * np.abs(f).max()
    path = ld("samples.bin").reshape(5, N)
    s = mg.MultigridMCSampler(op, par, nchains=1, seed=int(out["key"]))
    xs = np.zeros(N)
    for k in range(5):
        s.apply(f, xs)
        assert np.array_equal(xs, path[k]), k
    # fix_rhs: the same samples with f kept on the device
    assert np.array_equal(ld("samples_fixed.bin").reshape(5, N), path)
    # deep_copy with another generator: another key, another chain
    other = ld("samples_copy.bin").reshape(5, N)
    assert out["key_copy"] != out["key"] and not np.allclose(other, path)
    s2 = mg.MultigridMCSampler(op, par, nchains=1, seed=int(out["key_copy"]))
    xs = np.zeros(N)
    s2.apply(f, xs)
    assert np.array_equal(xs, other[0])
