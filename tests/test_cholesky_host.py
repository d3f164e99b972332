"""Full-problem Cholesky samplers (host code of libmgmc_amd.so; no GPU).

Pins:
- closed forms: with the normals replaced by the columns of the identity,
  the sampler's linear map M satisfies M M^T = Q^{-1} (Q = A + B Sigma^{-1}
  B^T assembled by scipy from the operators), and with xi = 0 the sample is
  the mean Q^{-1} f, for every factorisation, on 2D, 3D, uneven, variable-
  coefficient and FEM problems, with many overlapping measurement balls so
  that cliques straddle the nested-dissection separators;
- the reference's own sampler test (src/sampler/test_sampler.hh:56-153,
  :160-215): the 1D operator tridiag(-1, 6, -1) on 7 unknowns with and without
  B = 10 (e_3, e_4), Sigma = diag(4.2, 9.3), 5e5 samples, mean and covariance
  within 2e-3 in the maximum norm (the 1D lattice is the 2D lattice of 8 x 2
  cells, whose interior is one row of 7 vertices).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import multigridmc_amd as mg
from multigridmc_amd.cholesky import DenseCholeskySampler, LowRankCholeskySampler, SparseCholeskySampler

SAMPLERS = [SparseCholeskySampler, LowRankCholeskySampler, DenseCholeskySampler]


def dense_precision(op):
    A = sp.csr_matrix(op.get_sparse())
    if getattr(op, "m_lowrank", 0):
        B = sp.csr_matrix(op.get_B())
        A = A + B @ sp.diags(1.0 / op.get_Sigma()) @ B.T
    return A.toarray()


def _problems():
    from multigridmc_amd.operators import block_diffusion
    out = []
    lat = mg.Lattice(16, 16)
    c, y, v = mg.seeded_ball_measurements(lat, 12, 0.2, 1e-2, 3)
    out.append(("2d-16-balls", mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), c, 0.2, v)))
    lat = mg.Lattice(8, 8, 8)
    c, y, v = mg.seeded_ball_measurements(lat, 6, 0.25, 1e-2, 4)
    out.append(("3d-8-balls", mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), c, 0.25, v)))
    lat = mg.Lattice(13, 10)
    c, y, v = mg.seeded_ball_measurements(lat, 4, 0.15, 1e-3, 5)
    out.append(("2d-uneven", mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 3.0), c, 0.15, v)))
    lat = mg.Lattice(8, 8, 8)
    a = block_diffusion(lat, 2, 1.0, 10.0, 7)
    c, y, v = mg.seeded_ball_measurements(lat, 4, 0.2, 1e-2, 6)
    out.append(("3d-variable", mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0, diffusion=a), c, 0.2, v)))
    lat = mg.Lattice(12, 12)
    c, y, v = mg.seeded_ball_measurements(lat, 3, 0.2, 1e-2, 8)
    out.append(("2d-fem", mg.MeasuredOperator(mg.ShiftedLaplaceFEMOperator(lat, 10.0), c, 0.2, v)))
    out.append(("3d-prior", mg.ShiftedLaplaceFDOperator(mg.Lattice(8, 8, 8), 10.0)))
    return out


PROBLEMS = _problems()


@pytest.mark.parametrize("name,op", PROBLEMS, ids=[p[0] for p in PROBLEMS])
@pytest.mark.parametrize("cls", SAMPLERS, ids=lambda c: c.factorisation)
def test_closed_forms(name, op, cls):
    Q = dense_precision(op)
    N = Q.shape[0]
    Qi = np.linalg.inv(Q)
    s = cls(op, nchains=N, nthreads=4)
    X = s.sample(np.zeros(N), xi=np.eye(N))  # row c: M e_c
    assert np.abs(X.T @ X - Qi).max() <= 1e-12 * np.abs(Qi).max()
    f = np.random.default_rng(2).standard_normal(N)
    s1 = cls(op, nchains=1)
    mu = s1.sample(f, xi=np.zeros(N))[0]
    assert np.abs(mu - Qi @ f).max() <= 1e-12 * np.abs(Qi @ f).max()
    # the factor has at least the entries of the lower triangle of Q
    assert s.nnz_L >= (np.count_nonzero(np.tril(Q)) if cls is not LowRankCholeskySampler else 0)


@pytest.mark.parametrize("cls", SAMPLERS, ids=lambda c: c.factorisation)
def test_fix_rhs_threads_and_seed(cls):
    name, op = PROBLEMS[1]
    N = op.lattice.Nvertex
    f = np.random.default_rng(3).standard_normal(N)
    xi = np.random.default_rng(4).standard_normal((5, N))
    a = cls(op, nchains=5, nthreads=1)
    b = cls(op, nchains=5, nthreads=3)
    a.fix_rhs(f)
    x_fixed = a.sample(xi=xi)
    x_unfixed = b.sample(f, xi=xi)
    assert np.abs(x_fixed - x_unfixed).max() <= 1e-13 * np.abs(x_unfixed).max()
    a.unfix_rhs()
    with pytest.raises(mg.MGMCError):
        a.sample(None)
    # per-chain right-hand sides
    F = np.random.default_rng(5).standard_normal((5, N))
    x_pc = b.sample(F, xi=xi)
    for c in range(5):
        assert np.allclose(x_pc[c], b.sample(F[c], xi=xi)[c], rtol=0, atol=1e-13 * np.abs(x_pc).max())
    # Philox noise: same seed, same samples whatever the thread count; the
    # sample counter advances per call
    s1 = cls(op, nchains=9, nthreads=1, seed=77)
    s4 = cls(op, nchains=9, nthreads=4, seed=77)
    s1.fix_rhs(f)
    s4.fix_rhs(f)
    x1, x4 = s1.sample(), s4.sample()
    assert np.array_equal(x1, x4)
    assert not np.array_equal(x1, s1.sample())
    assert not np.array_equal(x1[0], x1[1])


def test_errors():
    op = PROBLEMS[0][1]
    with pytest.raises(mg.MGMCError):
        SparseCholeskySampler(op, nchains=0)
    big = mg.ShiftedLaplaceFDOperator(mg.Lattice(160, 160), 1.0)  # 25281 unknowns
    with pytest.raises(mg.MGMCError, match="16384"):
        DenseCholeskySampler(big)


class Operator1d:
    """tridiag(-1, 6, -1) on 7 unknowns (src/sampler/test_sampler.hh:56-84)."""

    def __init__(self, lowrank):
        self.lattice = mg.Lattice(8, 2)
        self.kappa_sq = 0.0
        n = self.lattice.Nvertex
        self.A = sp.diags([-np.ones(n - 1), 6.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
        self.m_lowrank = 2 if lowrank else 0
        self.B = sp.csr_matrix(([10.0, 10.0], ([3, 4], [0, 1])), shape=(n, 2))
        self.Sigma = np.array([4.2, 9.3])

    def get_sparse(self):
        return self.A

    def get_B(self):
        return self.B

    def get_Sigma(self):
        return self.Sigma


@pytest.mark.parametrize("lowrank", [False, True])
@pytest.mark.parametrize("cls", SAMPLERS, ids=lambda c: c.factorisation)
def test_reference_sampler_test_1d(cls, lowrank):
    op = Operator1d(lowrank)
    Q = dense_precision(op)
    n = Q.shape[0]
    mean_exact = np.random.default_rng(1342517).uniform(0.0, 1.0, n)
    f = Q @ mean_exact
    K, calls = 1000, 500
    s = cls(op, nchains=K, nthreads=4, seed=1342517)
    s.fix_rhs(f)
    X = np.concatenate([s.sample() for _ in range(calls)])
    Ex = X.mean(axis=0)
    cov = (X.T @ X) / X.shape[0] - np.outer(Ex, Ex)
    err_mean = np.abs(Ex - mean_exact).max()
    err_cov = np.abs(cov - np.linalg.inv(Q)).max()
    assert err_mean < 2e-3 and err_cov < 2e-3, (err_mean, err_cov)
