"""The reference's own visiting order on the GPU (mgmc_set_ordering with
MGMC_ORDER_LEXICOGRAPHIC, wavefront sweeps): single sweeps with and without
injected Gibbs noise, whole MG cycles, the loop solver's residual history
(src/solver/loop_solver.cc:24-42) and MGMC sampling statistics against the
oracle's lexicographic restatement (src/smoother/sor_smoother.cc:17-77,
src/sampler/sor_sampler.cc:37-59). FP64 at 1e-12; FP32 at 2e-6 relative to
the largest value (single precision, one sweep)."""
import numpy as np
import pytest

import multigridmc_amd as mg
from gpu_common import familywise_z, make_problem, moment_zscores, oracle_mg, params
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    (2, (16, 16), 0, 3),
    (2, (16, 32), 3, 3),
    (3, (8, 8, 8), 0, 2),
    (3, (16, 16, 16), 3, 3),
    (2, (24, 20), 3, 2),
    (3, (12, 16, 8), 3, 2),
]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("omega", [1.0, 0.8])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_lex_sweep_parity(case, omega, precision, tol):
    dim, n, nmeas, nlevel = case
    print(f"case: lexicographic sweep {case} omega {omega} {precision}", flush=True)
    op, ora = make_problem(dim, n, nmeas)
    p = params(nlevel, omega=omega)
    K = 2
    s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, ordering="lexicographic")
    omg = oracle_mg(ora, p, ordering=po.LEX)
    rng = np.random.default_rng(1)
    for level in range(nlevel):
        N = s.ndof(level)
        b, x = rng.standard_normal((K, N)), rng.standard_normal((K, N))
        xi, xl = rng.standard_normal((K, N)), rng.standard_normal((K, max(nmeas, 1)))
        for fw in (True, False):
            got = s.sweep(level, fw, False, b, x)
            for c in range(K):
                ref = omg.sor_smoother(level, fw, b[c], x[c])
                assert _rel(got[c], ref) < tol, (level, fw, c, _rel(got[c], ref))
            got = s.sweep(level, fw, True, b, x, xi=xi, xi_lr=xl[:, :nmeas] if nmeas else None)
            for c in range(K):
                ref = omg.sor_sampler_noise(level, fw, b[c], x[c], xi[c], xl[c, :nmeas] if nmeas else None)
                assert _rel(got[c], ref) < tol, (level, fw, c, _rel(got[c], ref))


@pytest.mark.parametrize("npre,npost,cycle,smoother", [(1, 1, 1, "SSOR"), (0, 1, 1, "SSOR"), (2, 1, 2, "SSOR"),
                                                        (1, 1, 1, "SOR"), (2, 2, 2, "SOR")])
def test_lex_mg_cycle_matches_oracle(npre, npost, cycle, smoother):
    op, ora = make_problem(3, (16, 16, 16), 3)
    p = params(3, smoother=smoother, cycle=cycle)
    p.npresmooth, p.npostsmooth = npre, npost
    pre = mg.MultigridPreconditioner(op, p, ordering="lexicographic")
    b = np.random.default_rng(5).standard_normal(op.lattice.Nvertex)
    got = pre.apply(b)
    ref = oracle_mg(ora, p, ordering=po.LEX).precondition(b)
    assert _rel(got, ref) < 1e-12


@pytest.mark.parametrize("dim,n,nmeas,nlevel", [(2, (64, 64), 0, 4), (2, (64, 64), 4, 4), (3, (32, 32, 32), 6, 4)])
@pytest.mark.parametrize("cycle", [1, 2])
def test_lex_mg_solve_same_residual_reduction(dim, n, nmeas, nlevel, cycle):
    """driver_mg's loop solver in the reference's order: the same iteration
    count, every residual norm equal to the oracle's to 1e-8 relative (so each
    per-cycle reduction rho_k is the reference's), and the solution within
    1e-10 of the oracle's. The multicolour solve reaches the same solution
    with its own (larger) rho_k, printed beside."""
    op, ora = make_problem(dim, n, nmeas, radius=0.1, var=1e-4)
    p = params(nlevel, cycle=cycle)
    b = np.random.default_rng(1482817).standard_normal(op.lattice.Nvertex)
    lex = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p, ordering="lexicographic"), rtol=1e-11, atol=1e-9,
                        maxiter=60)
    x = lex.apply(b)
    x_ref, hist_ref, it_ref = oracle_mg(ora, p, ordering=po.LEX).loop_solve(b, rtol=1e-11, atol=1e-9, maxiter=60)
    assert lex.niter == it_ref
    h = lex.history
    sig = hist_ref / hist_ref[0] > 1e-8
    assert np.allclose(h[sig], hist_ref[sig], rtol=1e-8)
    rho, rho_ref = h[1:] / h[:-1], hist_ref[1:] / hist_ref[:-1]
    s = sig[1:]
    assert np.all(np.abs(rho[s] / rho_ref[s] - 1.0) < 0.10)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-10
    mc = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p), rtol=1e-11, atol=1e-9, maxiter=60)
    x_mc = mc.apply(b)
    assert np.linalg.norm(x_mc - x_ref) / np.linalg.norm(x_ref) < 1e-10
    geo = lambda hh: float(np.exp(np.mean(np.log(hh[1:] / hh[:-1]))))  # noqa: E731
    print(f"mean rho: lexicographic GPU {geo(h):.4f}, reference {geo(hist_ref):.4f}, multicolour {geo(mc.history):.4f}")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lex_mgmc_posterior_moments(precision):
    """MGMC chains in the reference's order sample the exact posterior: probe
    means and covariances within the family-wise bound of their errors."""
    op, ora = make_problem(2, (16, 16), 2, radius=0.15, var=1e-3)
    N = op.lattice.Nvertex
    rng = np.random.default_rng(1342517)
    mean_exact = rng.uniform(size=N)
    Q = ora.dense_precision()
    f = Q @ mean_exact
    probes = rng.choice(N, size=8, replace=False)
    K = 128
    s = mg.MultigridMCSampler(op, params(3), nchains=K, precision=precision, seed=77, ordering="lexicographic")
    s.fix_rhs(f)
    s.set_state(np.zeros((K, N)))
    s.set_probes(probes)
    s.run(20, record=False)
    obs = s.run(600)
    cov = np.linalg.inv(Q)[np.ix_(probes, probes)]
    zm, zc = moment_zscores(obs, mean_exact[probes], cov)
    worst = max(np.abs(zm).max(), np.abs(zc).max())
    print(f"largest |z| {worst:.2f}, mean z^2 {np.mean(np.r_[zm, zc] ** 2):.2f}")
    assert worst < familywise_z(len(zm) + len(zc))
    assert np.mean(np.r_[zm, zc] ** 2) < 2.0
