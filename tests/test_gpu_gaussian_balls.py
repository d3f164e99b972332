"""Gaussian-weighted ball measurements (config 1's wording; the reference's
weight is the top-hat, src/linear_operator/measured_operator.hh:65): MGMC
chains of a 16^3 posterior with four Gaussian-weighted balls sample its
exact posterior (dense inverse), probe means and covariances within the
family-wise bound of their Monte Carlo errors, FP64 and FP32."""
import numpy as np
import pytest
import scipy.sparse as sp

import multigridmc_amd as mg
from gpu_common import familywise_z, moment_zscores

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gaussian_ball_posterior_sampling(precision):
    n = (16, 16, 16)
    lat = mg.Lattice(*n)
    rng = np.random.default_rng(31)
    centres = rng.uniform(0.25, 0.75, size=(4, 3))
    var = 1e-3 * (1 + rng.uniform(size=4))
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.2, var, profile="gaussian", width=0.4)
    N = lat.Nvertex
    B = op.get_B()
    Q = (op.get_sparse() + B @ sp.diags(1.0 / var) @ B.T).toarray()
    y = rng.standard_normal(4)
    f = B @ (y / var)
    cov_all = np.linalg.inv(Q)
    mean_all = cov_all @ f
    probes = rng.choice(N, size=10, replace=False)
    K = 64
    print(f"case: 16^3, 4 Gaussian balls, {precision}, {K} chains x 1500", flush=True)
    s = mg.MultigridMCSampler(op, mg.MultigridParameters(nlevel=3), nchains=K, precision=precision, seed=3)
    s.fix_rhs(f)
    s.set_state(np.zeros((K, N)))
    s.set_probes(probes)
    s.run(20, record=False)
    obs = s.run(1500)
    zm, zc = moment_zscores(obs, mean_all[probes], cov_all[np.ix_(probes, probes)])
    worst = max(np.abs(zm).max(), np.abs(zc).max())
    print(f"largest |z| {worst:.2f}, mean z^2 {np.mean(np.r_[zm, zc] ** 2):.2f}")
    assert worst < familywise_z(len(zm) + len(zc))
    assert np.mean(np.r_[zm, zc] ** 2) < 2.0
