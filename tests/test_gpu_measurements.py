"""Exact posterior statistics from multigrid solves and the batched
measurement harnesses (src/linear_operator/linear_operator.hh:131-189,
src/measurements/*.cc), on the GPU through the C ABI."""
import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import measurements as ms
from gpu_common import make_problem

pytestmark = pytest.mark.gpu


def _smw_reference(ora, xbar, y, b):
    """The reference's Sherman-Morrison-Woodbury forms with dense A^{-1}
    (linear_operator.hh:131-189), independent of the multigrid solve."""
    A = ora.csr(0).toarray()
    B = ora.csr(1).toarray()
    S = np.diag(ora.sigma())
    Ainv_B = np.linalg.solve(A, B)
    Ainv_b = np.linalg.solve(A, b)
    M = np.linalg.inv(S + B.T @ Ainv_B)
    x_post = xbar + Ainv_B @ M @ (y - B.T @ xbar)
    mean = b @ xbar + Ainv_b @ B @ M @ (y - B.T @ xbar)
    var = b @ Ainv_b - (B.T @ Ainv_b) @ M @ (B.T @ Ainv_b)
    return x_post, mean, var


@pytest.mark.parametrize("dim,n", [(3, (16, 16, 16)), (2, (32, 32))])
def test_posterior_mean_and_observed_moments_match_smw(dim, n):
    op, ora = make_problem(dim, n, 4)
    rng = np.random.default_rng(5)
    N = op.lattice.Nvertex
    xbar = rng.standard_normal(N) * 0.1
    y = rng.standard_normal(op.m_lowrank)
    x_post = ms.posterior_mean(op, xbar, y)
    obs = [mg.measurement_vector(op.lattice, [0.4] * dim, 0.2), mg.measurement_vector(op.lattice, [0.7] * dim, 0.0)]
    means, cov = ms.observed_moments(op, xbar, y, obs)
    for p, b in enumerate(obs):
        bd = np.zeros(N)
        bd[b[0]] = b[1]
        xr, mr, vr = _smw_reference(ora, xbar, y, bd)
        assert np.abs(x_post - xr).max() < 1e-9 * np.abs(xr).max()
        assert abs(means[p] - mr) < 1e-9 * max(abs(mr), 1e-3)
        assert abs(cov[p, p] - vr) < 1e-9 * vr
    # off-diagonal against the dense posterior covariance
    Bo = np.stack([ms._dense(op, b) for b in obs])
    C = Bo @ np.linalg.solve(ora.dense_precision(), Bo.T)
    assert np.abs(cov - C).max() < 1e-9 * np.abs(C).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_observables_equal_dot_products(precision):
    op, _ = make_problem(3, (16, 16, 16), 3)
    N = op.lattice.Nvertex
    K = 3
    s = mg.MultigridMCSampler(op, mg.MultigridParameters(nlevel=2), nchains=K, precision=precision, seed=3)
    f = np.random.default_rng(1).standard_normal(N)
    s.fix_rhs(f)
    s.set_state(np.zeros((K, N)))
    obs = [mg.measurement_vector(op.lattice, [0.3, 0.5, 0.6], 0.25), mg.measurement_vector(op.lattice, [0.5] * 3, 0.0),
           (np.array([0, N - 1]), np.array([2.0, -1.0]))]
    s.set_observables(obs)
    z = s.run(4, use_graph=False)
    x = s.get_state()
    Bo = np.stack([ms._dense(op, b) for b in obs])
    want = x @ Bo.T
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.abs(z[-1] - want).max() <= tol * np.abs(want).max()
    # probes are observables with one entry of weight 1
    s.set_probes(np.array([5, 17]))
    z2 = s.run(1, use_graph=True)
    assert np.abs(z2[0] - s.get_state()[:, [5, 17]]).max() <= tol * np.abs(z2).max()


def test_measurement_harnesses(tmp_path):
    op, _ = make_problem(2, (32, 32), 2)
    rng = np.random.default_rng(9)
    y = rng.standard_normal(op.m_lowrank)
    b = mg.measurement_vector(op.lattice, [0.45, 0.55], 0.1)
    xbar = np.zeros(op.lattice.Nvertex)
    x_post = ms.posterior_mean(op, xbar, y)
    mean_exact, var_exact = ms.observed_mean_and_variance(op, xbar, y, b)
    exact = (x_post, mean_exact, var_exact)
    s = mg.MultigridMCSampler(op, mg.MultigridParameters(nlevel=3), nchains=128, seed=11)
    # convergence of the first two moments from x = 0
    fn = tmp_path / "convergence.txt"
    dm, em, dv, ev = ms.measure_convergence(s, 512, 6, b, y, filename=str(fn), exact=exact)
    text = fn.read_text().splitlines()
    assert text[0].startswith("**** q_k = |E[z^k] - E[z]|")
    assert sum(1 for ln in text if ln.strip().startswith("mean")) == 7
    assert sum(1 for ln in text if ln.strip().startswith("variance")) == 7
    assert dm[0] == pytest.approx(abs(mean_exact), rel=1e-12) and dv[0] == pytest.approx(var_exact, rel=1e-12)
    # after a few MGMC steps the chains are stationary within the statistical errors
    assert np.all(dm[3:] < 5 * em[3:] + 1e-12) and np.all(dv[3:] < 5 * ev[3:] + 1e-12)
    # mean squared error of the running average decreases
    fn2 = tmp_path / "mse.txt"
    mse, err = ms.measure_mean_squared_error(s, 256, 8, b, y, filename=str(fn2), exact=exact)
    assert len(fn2.read_text().splitlines()) == 8
    assert mse[-1] < 0.5 * mse[1]
    # timed sampling with the exact moments
    ms_per, data, (me, ve) = ms.measure_sampling_time(s, 10, 200, b, y, "MGMC", filename=str(tmp_path / "ts.txt"),
                                                      exact=exact, verbose=False)
    assert data.shape == (200, 128) and ms_per > 0
    tau = mg.tau_int_chains(data)[0]
    se = np.sqrt(ve * tau / data.size)
    assert abs(data.mean() - me) < 4 * se
    assert abs(data.var() / ve - 1) < 0.05
