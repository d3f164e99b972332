"""The benchmark configurations at full size, checked statistically against
the exact posterior (the test of src/sampler/test_sampler.hh:113-153 and the
closed forms of src/linear_operator/linear_operator.hh:131-189):

* 2D 64 x 64 prior (63^2 unknowns), FP64, one chain, 10,000 samples, 4
  levels, against the dense inverse of the oracle's precision matrix;
* 128^3 posterior with 32 balls (radius 0.05, variance 1e-4), FP64, one
  chain, 20,000 samples, 5 levels;
* the same posterior as 64 FP32 chains x 2,000 samples, with the probe
  covariance error against the FP64 exact posterior stated and bounded.

Every probe mean and covariance entry must lie within the family-wise
(Bonferroni, 1 %) bound of its Monte Carlo standard error - that is, the
north star's 3 standard errors per quantity corrected for the number of
quantities tested - and the mean square of the standardised deviations must
be consistent with pure Monte Carlo noise. The exact 128^3 moments come from
FP64 multigrid solves (solution tolerance 1e-12)."""
import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import measurements
from oracle import pyoracle as po

from gpu_common import familywise_z, moment_zscores

pytestmark = pytest.mark.gpu

V11 = mg.MultigridParameters(nlevel=5, smoother="SSOR", npresmooth=1, npostsmooth=1, omega=1.0, cycle=1,
                             coarse_solver="Cholesky")


def _check(label, zm, zc):
    q = len(zm) + len(zc)
    zmax = familywise_z(q)
    worst = max(np.abs(zm).max(), np.abs(zc).max())
    msq = float(np.mean(np.concatenate([zm, zc]) ** 2))
    print(f"{label}: {q} quantities, largest |z| {worst:.2f} (bound {zmax:.2f}), mean z^2 {msq:.2f}", flush=True)
    assert worst < zmax
    assert msq < 2.0


def _sample(op, f, params, nchains, precision, probes, nwarm, nsamples, seed):
    s = mg.MultigridMCSampler(op, params, nchains=nchains, precision=precision, seed=seed)
    s.fix_rhs(f)
    s.set_state(np.zeros((nchains, op.lattice.Nvertex)))
    s.set_probes(probes)
    s.run(nwarm, record=False)
    return s.run(nsamples)


def test_config0_2d_prior_fp64_single_chain():
    n = (64, 64)
    op = mg.ShiftedLaplaceFDOperator(mg.Lattice(*n), 10.0)
    N = op.lattice.Nvertex
    probes = np.sort(np.random.default_rng(160).choice(N, size=16, replace=False))
    p = mg.MultigridParameters(nlevel=4, smoother="SSOR", npresmooth=1, npostsmooth=1, omega=1.0, cycle=1,
                               coarse_solver="Cholesky")
    print("case: config 0, 2D 64^2 prior, FP64, 1 chain x 10,000", flush=True)
    obs = _sample(op, np.zeros(N), p, 1, "fp64", probes, 100, 10000, 11)
    cov = np.linalg.inv(po.Operator.fd(2, n, 10.0).dense_precision())[np.ix_(probes, probes)]
    _check("config 0", *moment_zscores(obs, np.zeros(len(probes)), cov))


@pytest.fixture(scope="module")
def posterior_128():
    n, m = 128, 32
    lat = mg.Lattice(n, n, n)
    centres, y, var = mg.seeded_ball_measurements(lat, m, 0.05, 1e-4, 20240)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.05, var)
    f = op.get_B() @ (y / op.get_Sigma())  # Q x_post for a zero prior mean
    probes = np.sort(np.random.default_rng(20241).choice(lat.Nvertex, size=32, replace=False))
    obsv = [(np.array([q]), np.array([1.0])) for q in probes]
    mean, cov = measurements.observed_moments(op, np.zeros(lat.Nvertex), y, obsv)
    return op, f, probes, mean, cov


def test_config1_128cube_posterior_fp64_single_chain(posterior_128):
    op, f, probes, mean, cov = posterior_128
    print("case: config 1, 128^3 posterior, FP64, 1 chain x 20,000", flush=True)
    obs = _sample(op, f, V11, 1, "fp64", probes, 50, 20000, 21)
    _check("config 1", *moment_zscores(obs, mean, cov))


def test_config2_128cube_posterior_fp32_64_chains(posterior_128):
    op, f, probes, mean, cov = posterior_128
    print("case: config 2, 128^3 posterior, FP32, 64 chains x 2,000", flush=True)
    obs = _sample(op, f, V11, 64, "fp32", probes, 20, 2000, 31)
    _check("config 2", *moment_zscores(obs, mean, cov))
    # the north star's covariance error against the FP64 exact posterior
    z = obs.reshape(-1, len(probes))
    n = z.shape[0]
    tau = max(mg.tau_int_chains(obs[:, :, p], mean=mean[p])[0] for p in range(len(probes)))
    d = z - z.mean(axis=0)
    C = d.T @ d / n
    rel_var = (np.diag(C) - np.diag(cov)) / np.diag(cov)
    rms = float(np.sqrt(np.mean(rel_var ** 2)))
    noise = float(np.sqrt(2 * tau / n))
    frob = float(np.linalg.norm(C - cov) / np.linalg.norm(cov))
    dg = np.diag(cov)
    floor = float(np.sqrt(np.sum(np.outer(dg, dg) + cov ** 2) * tau / n) / np.linalg.norm(cov))
    print(f"config 2 covariance error vs FP64 exact: RMS relative variance error {rms:.4f} (Monte Carlo noise "
          f"{noise:.4f}), relative Frobenius {frob:.4f} (noise floor {floor:.4f}), tau {tau:.3f}", flush=True)
    assert rms <= 0.01
    assert frob <= 0.01 + 2 * floor
