"""Variable coefficients on the GPU (per-vertex stencils on every level:
k_sweep_var, k_residual_var, per-vertex Galerkin coarse operators) against
the oracle given the same operator as a CSR matrix (its multicolour and
lexicographic restatements of src/smoother/sor_smoother.cc,
src/sampler/sor_sampler.cc, src/sampler/multigridmc_sampler.cc), for a
cell-wise diffusion coefficient -div(a grad) + kappa^2 and for the
reference's periodic correlation-length model kappa^2(x)
(src/linear_operator/correlationlength_model.hh:83-105)."""
import numpy as np
import pytest
import scipy.sparse as sp

import multigridmc_amd as mg
from gpu_common import familywise_z, moment_zscores, oracle_mg, params
from multigridmc_amd.operators import periodic_kappa_sq
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    # dim, n, nmeas, nlevel
    (2, (16, 16), 0, 3),
    (2, (16, 32), 3, 3),
    (3, (8, 8, 8), 0, 2),
    (3, (16, 16, 16), 3, 3),
    (2, (24, 20), 3, 2),
    (3, (12, 16, 8), 3, 2),
]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def variable_problem(dim, n, nmeas, seed=17, radius=0.15, var=1e-3, kappa_field=False):
    lat = mg.Lattice(*n)
    rng = np.random.default_rng(seed)
    a = rng.uniform(1.0, 10.0, size=int(np.prod(n)))
    kap = 10.0 + 5.0 * rng.uniform(size=lat.Nvertex) if kappa_field else 10.0
    prior = mg.ShiftedLaplaceFDOperator(lat, kap, diffusion=a)
    A = prior.get_sparse().tocoo()
    ora = po.Operator.custom(dim, n, A.row, A.col, A.data)
    if nmeas:
        centres = rng.uniform(0.2, 0.8, size=(nmeas, dim))
        variance = var * (1.0 + rng.uniform(size=nmeas))
        op = mg.MeasuredOperator(prior, centres, radius, variance)
        ora.add_measurements(centres, radius, variance)
    else:
        op = prior
    return op, ora


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("ordering", ["multicolour", "lexicographic"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 3e-6)])
def test_variable_sweep_parity(case, ordering, precision, tol):
    dim, n, nmeas, nlevel = case
    print(f"case: variable sweep {case} {ordering} {precision}", flush=True)
    op, ora = variable_problem(dim, n, nmeas, kappa_field=True)
    p = params(nlevel, omega=0.9)
    K = 2
    s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, ordering=ordering)
    omg = oracle_mg(ora, p, ordering=po.LEX if ordering == "lexicographic" else po.MULTICOLOUR)
    rng = np.random.default_rng(1)
    for level in range(nlevel):
        N = s.ndof(level)
        b, x = rng.standard_normal((K, N)), rng.standard_normal((K, N))
        xi, xl = rng.standard_normal((K, N)), rng.standard_normal((K, max(nmeas, 1)))
        for fw in (True, False):
            got = s.sweep(level, fw, False, b, x)
            for c in range(K):
                assert _rel(got[c], omg.sor_smoother(level, fw, b[c], x[c])) < tol, (level, fw, c)
            got = s.sweep(level, fw, True, b, x, xi=xi, xi_lr=xl[:, :nmeas] if nmeas else None)
            for c in range(K):
                ref = omg.sor_sampler_noise(level, fw, b[c], x[c], xi[c], xl[c, :nmeas] if nmeas else None)
                assert _rel(got[c], ref) < tol, (level, fw, c)


@pytest.mark.parametrize("case", CASES)
def test_variable_transfer_and_coarse_parity(case):
    dim, n, nmeas, nlevel = case
    op, ora = variable_problem(dim, n, nmeas)
    p = params(nlevel)
    K = 2
    s = mg.MultigridMCSampler(op, p, nchains=K)
    omg = oracle_mg(ora, p)
    rng = np.random.default_rng(2)
    for level in range(nlevel - 1):
        N, Nc = s.ndof(level), s.ndof(level + 1)
        f, x = rng.standard_normal((K, N)), rng.standard_normal((K, N))
        got = s.residual_restrict(level, f, x)
        for c in range(K):
            assert _rel(got[c], omg.restrict(level, f[c] - omg.level_apply(level, x[c]))) < 1e-12
    Nc = s.ndof(nlevel - 1)
    f, xi = rng.standard_normal((K, Nc)), rng.standard_normal((K, Nc))
    got, got_d = s.coarse(f, xi_mode=2, xi=xi), s.coarse(f, xi_mode=0)
    for c in range(K):
        assert _rel(got[c], omg.coarse_sample_noise(f[c], xi[c])) < 1e-11
        assert _rel(got_d[c], omg.coarse_solve(f[c])) < 1e-11
    x = rng.standard_normal((K, op.lattice.Nvertex))
    y = s.apply_operator(x)
    for c in range(K):
        assert _rel(y[c], ora.apply(x[c])) < 1e-13


@pytest.mark.parametrize("cycle", [1, 2])
def test_variable_mg_cycle_and_solve(cycle):
    op, ora = variable_problem(3, (32, 32, 32), 4, radius=0.1, var=1e-4)
    p = params(4, cycle=cycle)
    b = np.random.default_rng(1482817).standard_normal(op.lattice.Nvertex)
    pre = mg.MultigridPreconditioner(op, p)
    omg = oracle_mg(ora, p)
    assert _rel(pre.apply(b), omg.precondition(b)) < 1e-12
    solver = mg.LoopSolver(op, pre, rtol=1e-11, atol=1e-9, maxiter=80)
    x = solver.apply(b)
    x_ref, hist, it = omg.loop_solve(b, rtol=1e-11, atol=1e-9, maxiter=80)
    assert solver.niter == it
    sig = hist / hist[0] > 1e-8
    assert np.allclose(solver.history[sig], hist[sig], rtol=1e-8)
    Q = (op.get_sparse() + op.get_B() @ sp.diags(1.0 / op.get_Sigma()) @ op.get_B().T).tocsc()
    x_dense = sp.linalg.spsolve(Q, b)
    assert np.linalg.norm(x - x_dense) / np.linalg.norm(x_dense) < 1e-10
    # FP32 cycles with FP64 refinement reach the same solution
    x32 = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p, precision="fp32"), rtol=1e-11, atol=1e-9,
                        maxiter=80).apply(b)
    assert np.linalg.norm(x32 - x_dense) / np.linalg.norm(x_dense) < 1e-10


def test_periodic_correlation_length_model_solves_like_the_reference():
    """The reference's periodic kappa^2(x) in its own (lexicographic) order:
    the loop solver's residual history and solution equal the oracle's
    restatement of the reference operator (fd_periodic)."""
    n = (32, 32, 32)
    lat = mg.Lattice(*n)
    op = mg.ShiftedLaplaceFDOperator(lat, periodic_kappa_sq(lat, 0.1, 0.3))
    ora = po.Operator.fd_periodic(3, n, 0.1, 0.3)
    p = params(4)
    b = np.random.default_rng(1482817).standard_normal(lat.Nvertex)
    solver = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p, ordering="lexicographic"), rtol=1e-11, atol=1e-9,
                           maxiter=60)
    x = solver.apply(b)
    x_ref, hist, it = oracle_mg(ora, p, ordering=po.LEX).loop_solve(b, rtol=1e-11, atol=1e-9, maxiter=60)
    assert solver.niter == it
    sig = hist / hist[0] > 1e-8
    assert np.allclose(solver.history[sig], hist[sig], rtol=1e-8)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_variable_posterior_sampling(precision):
    op, ora = variable_problem(3, (16, 16, 16), 4, radius=0.2)
    N = op.lattice.Nvertex
    Q = ora.dense_precision()
    rng = np.random.default_rng(9)
    mean_exact = rng.standard_normal(N)
    f = Q @ mean_exact
    probes = rng.choice(N, size=10, replace=False)
    K = 64
    s = mg.MultigridMCSampler(op, params(3), nchains=K, precision=precision, seed=8)
    s.fix_rhs(f)
    s.set_state(np.zeros((K, N)))
    s.set_probes(probes)
    s.run(20, record=False)
    obs = s.run(1500)
    cov = np.linalg.inv(Q)[np.ix_(probes, probes)]
    zm, zc = moment_zscores(obs, mean_exact[probes], cov)
    worst = max(np.abs(zm).max(), np.abs(zc).max())
    print(f"largest |z| {worst:.2f}, mean z^2 {np.mean(np.r_[zm, zc] ** 2):.2f}")
    assert worst < familywise_z(len(zm) + len(zc))
    assert np.mean(np.r_[zm, zc] ** 2) < 2.0
