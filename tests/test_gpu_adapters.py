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


def test_level_component_adapters_through_adapter_demo(tmp_path):
    """The Smoother, SORSampler / SSORSampler and IntergridOperator shapes
    (src/smoother/smoother.hh:29, sor_smoother.cc:41-77, ssor_smoother.cc:9-15,
    src/sampler/sor_sampler.cc:37-59, src/intergrid/intergrid_operator.hh:74-120)
    against the oracle (multicolour twin, omega 0.8): SOR in both directions and
    SSOR with nsmooth 2 on the posterior, the reference's nsmooth = 2 SOR
    smoother (twice two sparse sweeps) on the prior, the SOR sampler's first
    two calls given the normals it drew (fresh on each call), the SSOR
    sampler bit for bit against the Python SSORSampler with its key, and the
    restriction and prolongation."""
    print("case: adapter_demo level components, 2D 32^2", flush=True)
    p = subprocess.run([os.path.join(BIN, "adapter_demo"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    out = dict(ln.split(" ", 1) for ln in p.stdout.splitlines() if ln.startswith("key"))
    ld = lambda name: np.fromfile(tmp_path / name, dtype=np.float64)  # noqa: E731
    centres = np.array([[0.3, 0.4], [0.7, 0.55], [0.45, 0.75]])
    var = np.array([1e-3, 2e-3, 1.5e-3])
    post = po.Operator.fd(2, (32, 32), 10.0)
    post.add_measurements(centres, 0.1, var)
    prior = po.Operator.fd(2, (32, 32), 10.0)
    b, x0 = ld("b.bin"), ld("x0.bin")
    mgp = po.Multigrid(post, 1, omega=0.8, ordering=po.MULTICOLOUR)
    mgq = po.Multigrid(prior, 1, omega=0.8, ordering=po.MULTICOLOUR)

    def close(a, ref, tol=1e-12):
        err = np.abs(a - ref).max() / np.abs(ref).max()
        assert err < tol, err

    close(ld("sor_fwd.bin"), mgp.sor_smoother(0, True, b, x0))
    close(ld("sor_bwd.bin"), mgp.sor_smoother(0, False, b, x0))
    ref = x0
    for _ in range(2):
        ref = mgp.sor_smoother(0, False, b, mgp.sor_smoother(0, True, b, ref))
    close(ld("ssor2.bin"), ref)
    ref = x0
    for _ in range(4):  # without measurements the smoother is the sparse sweep
        ref = mgq.sor_smoother(0, True, b, ref)
    close(ld("sor2_prior.bin"), ref)
    xi1, xi2 = ld("sor_sampler_xi1.bin"), ld("sor_sampler_xi2.bin")
    assert not np.allclose(xi1, xi2)
    x1 = mgq.sor_sampler_noise(0, True, b, x0, xi1, None)
    close(ld("sor_sampler_x1.bin"), x1)
    close(ld("sor_sampler_x2.bin"), mgq.sor_sampler_noise(0, True, b, x1, xi2, None))
    lat = mg.Lattice(32, 32)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.1, var)
    s = mg.SSORSampler(op, omega=1.0, nsmooth=1, seed=int(out["key_ssor"]))
    xs = np.zeros(lat.Nvertex)
    path = ld("ssor_sampler.bin").reshape(3, -1)
    for k in range(3):
        s.apply(b, xs)
        assert np.array_equal(xs, path[k]), k
    xc = ld("restrict.bin")
    close(xc, post.restrict(x0), 1e-14)
    close(ld("prolongate.bin"), post.prolongate_add(0.7, xc, x0), 1e-14)
