"""Pins the CPU oracle's smoothers, solvers and samplers.

Checks that do not re-use the oracle's formulas:
* the reference's invariance and convergence tests (test_smoother.hh,
  test_solver.hh) on the FD prior,
* dense brute-force splittings M = T + B Sigma^-1 B^T for one SOR sweep,
* exactness of every Gibbs sweep and of a whole MGMC cycle, shown by building
  the affine map x' = Phi x + c + Psi xi from injected noise and checking that
  N(Q^-1 f, Q^-1) is left invariant to rounding,
* the reference's statistical sampler test on its 1D TestOperator.
"""
import numpy as np
import pytest


def _posterior(oracle, dim, n, nmeas, seed=1212417, radius=0.1, var_scale=1e-6):
    op = oracle.Operator.fd(dim, n, 10.0)
    if nmeas:
        rng = np.random.default_rng(seed)
        centres = rng.uniform(0.15, 0.85, size=(nmeas, dim))
        var = var_scale * (1.0 + 2.0 * rng.uniform(size=nmeas))
        op.add_measurements(centres, radius, var)
    return op


def _dense(op):
    A = op.csr(0).toarray()
    if op.m:
        B = op.csr(1).toarray()
        return A, B, np.diag(op.sigma())
    return A, np.zeros((op.N, 0)), np.zeros((0, 0))


def _order(op, ordering, forward, n, dim):
    """Visiting order: lexicographic, or colour-major multicolour."""
    N = op.N
    if ordering == 0:
        o = np.arange(N)
    else:
        idx = np.stack(np.unravel_index(np.arange(N), [v - 1 for v in n][::-1]), axis=1)[:, ::-1] + 1
        A = op.csr(0)
        coo = A.tocoo()
        off = np.abs(idx[coo.row] - idx[coo.col])
        star = np.all(np.count_nonzero(off, axis=1) <= 1)
        if star:
            colour = idx.sum(axis=1) & 1
        else:
            colour = sum((idx[:, d] & 1) << d for d in range(dim))
        o = np.concatenate([np.where(colour == c)[0] for c in range(colour.max() + 1)])
    return o if forward else o[::-1]


@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("lowrank", [False, True])
def test_ssor_smoother_invariance(oracle, ordering, lowrank):
    """test_smoother.hh:90-114: SSOR (omega = 0.8) leaves the exact solution fixed."""
    op = _posterior(oracle, 2, (32, 32), 10 if lowrank else 0)
    mg = oracle.Multigrid(op, 1, omega=0.8, ordering=ordering)
    x_exact = np.random.default_rng(1212417).standard_normal(op.N)
    b = op.apply(x_exact)
    x = mg.sor_smoother(0, True, b, x_exact)
    x = mg.sor_smoother(0, False, b, x)
    assert np.linalg.norm(x - x_exact) / np.linalg.norm(x_exact) < 1e-12


@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("forward", [True, False])
def test_sor_sweep_is_splitting(oracle, ordering, forward):
    """x' = M^-1 (b + N x) with M = T + B Sigma^-1 B^T, T = D/omega + strict
    lower part in visiting order (sor_smoother.hh:22-42)."""
    n = (8, 8, 8)
    op = _posterior(oracle, 3, n, 3, radius=0.2, var_scale=1e-3)
    omega = 0.8
    mg = oracle.Multigrid(op, 1, omega=omega, ordering=ordering)
    A, B, S = _dense(op)
    o = _order(op, ordering, forward, n, 3)
    pos = np.empty_like(o)
    pos[o] = np.arange(len(o))
    T = np.where(pos[None, :] < pos[:, None], A, 0.0) + np.diag(np.diag(A)) / omega
    M = T + B @ np.linalg.solve(S, B.T)
    Q = A + B @ np.linalg.solve(S, B.T)
    Nm = M - Q
    rng = np.random.default_rng(5)
    x, b = rng.standard_normal(op.N), rng.standard_normal(op.N)
    out = mg.sor_smoother(0, forward, b, x)
    ref = np.linalg.solve(M, b + Nm @ x)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-12


def _affine_map(apply, nx, nxi):
    """Columns of x' = Phi x + c + Psi xi from unit vectors."""
    c = apply(np.zeros(nx), np.zeros(nxi))
    Phi = np.stack([apply(e, np.zeros(nxi)) - c for e in np.eye(nx)], axis=1)
    Psi = np.stack([apply(np.zeros(nx), e) - c for e in np.eye(nxi)], axis=1)
    return Phi, c, Psi


@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("forward", [True, False])
@pytest.mark.parametrize("omega", [1.0, 0.7])
def test_gibbs_sweep_exact(oracle, ordering, forward, omega):
    """One SOR Gibbs sweep with SMW correction (sor_sampler.cc:37-59) leaves
    N(Q^-1 f, Q^-1) invariant."""
    n = (8, 8)
    op = _posterior(oracle, 2, n, 2, radius=0.2, var_scale=1e-3)
    mg = oracle.Multigrid(op, 1, omega=omega, ordering=ordering)
    N, m = op.N, op.m
    Q = op.dense_precision()
    f = np.random.default_rng(2).standard_normal(N)
    Phi, c, Psi = _affine_map(lambda x, xi: mg.sor_sampler_noise(0, forward, f, x, xi[:N], xi[N:]), N, N + m)
    Cov = np.linalg.inv(Q)
    mu = Cov @ f
    assert np.allclose(Phi @ mu + c, mu, rtol=0, atol=1e-11 * np.abs(mu).max())
    assert np.allclose(Phi @ Cov @ Phi.T + Psi @ Psi.T, Cov, rtol=0, atol=1e-11 * np.abs(Cov).max())


@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("smoother,cycle", [(1, 1), (0, 1), (1, 2)])
def test_mgmc_cycle_exact(oracle, ordering, smoother, cycle):
    """A whole MGMC cycle (multigridmc_sampler.cc:109-145) leaves the posterior
    invariant: stationary mean Q^-1 f and covariance Q^-1, to rounding."""
    n = (16, 16)
    op = _posterior(oracle, 2, n, 2, radius=0.15, var_scale=1e-3)
    mg = oracle.Multigrid(op, 3, smoother=smoother, cycle=cycle, ordering=ordering)
    N = op.N
    nxi = mg.mgmc_noise_count()
    Q = op.dense_precision()
    f = np.random.default_rng(4).standard_normal(N)
    Phi, c, Psi = _affine_map(lambda x, xi: mg.mgmc_apply_noise(f, x, xi), N, nxi)
    Cov = np.linalg.inv(Q)
    mu = Cov @ f
    assert np.allclose(Phi @ mu + c, mu, rtol=0, atol=1e-10 * np.abs(mu).max())
    assert np.allclose(Phi @ Cov @ Phi.T + Psi @ Psi.T, Cov, rtol=0, atol=1e-10 * np.abs(Cov).max())
    # a good multigrid sampler contracts fast: spectral radius of Phi well below 1
    assert np.max(np.abs(np.linalg.eigvals(Phi))) < 0.5


def test_coarse_cholesky_sample(oracle):
    """x = L^-T (xi + L^-1 f) (cholesky_sampler.hh:52-66) with Q = L L^T."""
    op = _posterior(oracle, 3, (16, 16, 16), 4, radius=0.2, var_scale=1e-3)
    mg = oracle.Multigrid(op, 3)
    Nc = mg.ndof(2)
    Qc = mg.level_csr(2, 0).toarray()
    Bc = mg.level_csr(2, 1).toarray()
    Qc = Qc + Bc @ np.diag(1 / op.sigma()) @ Bc.T
    L = np.linalg.cholesky(Qc)
    rng = np.random.default_rng(9)
    f, xi = rng.standard_normal(Nc), rng.standard_normal(Nc)
    x = mg.coarse_sample_noise(f, xi)
    ref = np.linalg.solve(L.T, xi + np.linalg.solve(L, f))
    assert np.allclose(x, ref, rtol=1e-12, atol=1e-14)
    assert np.allclose(mg.coarse_solve(f), np.linalg.solve(Qc, f), rtol=1e-11)


@pytest.mark.parametrize("dim,n,nlevel,nmeas", [(2, (64, 64), 4, 0), (2, (64, 64), 4, 6), (3, (16, 16, 16), 3, 4)])
@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("cycle", [1, 2])
def test_mg_solve(oracle, dim, n, nlevel, nmeas, ordering, cycle):
    """test_solver.hh:113-175: V(1,1) SSOR MG + LoopSolver to 1e-10 relative error."""
    op = _posterior(oracle, dim, n, nmeas, radius=0.1, var_scale=1e-4)
    mg = oracle.Multigrid(op, nlevel, cycle=cycle, ordering=ordering)
    b = np.random.default_rng(1482817).standard_normal(op.N)
    x, hist, it = mg.loop_solve(b, rtol=1e-12, atol=1e-9, maxiter=60)
    assert it > 0
    x_exact = op.dense_solve(b) if op.N < 5000 else None
    if x_exact is not None:
        assert np.linalg.norm(x - x_exact) / np.linalg.norm(x_exact) < 1e-10
    rho = hist[1:] / hist[:-1]
    assert rho.max() < 0.35


def _test_operator_1d(oracle, lowrank):
    """TestOperator1d (test_sampler.hh:47-88): tridiag(-1, 6, -1), 7 unknowns."""
    rows, cols, vals = [], [], []
    for i in range(7):
        rows.append(i); cols.append(i); vals.append(6.0)
        if i > 0:
            rows.append(i); cols.append(i - 1); vals.append(-1.0)
        if i < 6:
            rows.append(i); cols.append(i + 1); vals.append(-1.0)
    op = oracle.Operator.custom(1, (8,), rows, cols, vals)
    if lowrank:
        op.set_lowrank(2, [3, 4], [0, 1], [10.0, 10.0], [4.2, 9.3])
    return op


@pytest.mark.parametrize("lowrank", [False, True])
def test_mgmc_1d_statistics(oracle, lowrank):
    """test_sampler.hh:113-153,237-264: MGMC (3 levels, SSOR, Cholesky) on the
    1D TestOperator; mean and covariance in the L-infinity norm. The reference
    draws 1e6 samples at tolerance 2e-3; here 2e5 samples at 5e-3."""
    op = _test_operator_1d(oracle, lowrank)
    mg = oracle.Multigrid(op, 3)
    mean_exact = np.random.default_rng(1342517).uniform(size=op.N)
    Q = op.dense_precision()
    f = Q @ mean_exact
    x0 = np.zeros((1, op.N))
    x0, _, _ = mg.chains(f, x0, 1000)
    probes = np.arange(op.N)
    _, obs, _ = mg.chains(f, x0, 200000, probes=probes)
    s = obs[:, 0, :]
    mean = s.mean(axis=0)
    cov = np.cov(s.T, bias=True)
    assert np.abs(mean - mean_exact).max() < 5e-3
    assert np.abs(cov - np.linalg.inv(Q)).max() < 5e-3
