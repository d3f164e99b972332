"""Exact posterior statistics from FP64 multigrid solves, and the reference's
measurement harnesses run on batched chains.

`posterior_mean` and `observed_mean_and_variance` give the quantities of
LinearOperator::mean and LinearOperator::observed_mean_and_variance
(src/linear_operator/linear_operator.hh:131-189), which the reference gets
from a sparse Cholesky factorisation of the fine prior precision. Here they
come from the multigrid loop solver on the GPU (FP64 cycle, FP64 residuals):

    x_post = Q^{-1} (A xbar + B Sigma^{-1} y),   Q = A + B Sigma^{-1} B^T,
    E[b^T x] = b^T x_post,   Var[b^T x] = b^T Q^{-1} b,

which equal the reference's Sherman-Morrison-Woodbury forms. Several
observation vectors are solved together as the chains of one handle.

`measure_sampling_time`, `measure_convergence` and `measure_mean_squared_error`
follow src/measurements/{sampling_time,convergence,mean_squared_error}.cc: the
same observable z = b^T x, the same estimators and the same output files; the
independent chains the reference runs on OpenMP threads (one deep-copied
sampler per thread) are the K chains of one batched sampler here.
"""
import time

import numpy as np

from .sampler import LoopSolver, MultigridParameters, MultigridPreconditioner


def _solve(op, rhs, params=None, rtol=1e-12, atol=1e-14, maxiter=200):
    """Q^{-1} rhs for each row of rhs (FP64 multigrid loop solver)."""
    rhs = np.atleast_2d(np.asarray(rhs, dtype=np.float64))
    pre = MultigridPreconditioner(op, params or MultigridParameters(nlevel=_nlevel(op)), precision="fp64",
                                  nchains=rhs.shape[0])
    solver = LoopSolver(op, pre, rtol=rtol, atol=atol, maxiter=maxiter)
    out = solver.apply(rhs)
    return np.asarray(out).reshape(rhs.shape), solver


def _nlevel(op):
    """Levels down to a coarsest lattice with at most 8 cells per direction."""
    n = min(op.lattice.n)
    L = 1
    while n % 2 == 0 and n // 2 >= 2 and n > 8:
        n //= 2
        L += 1
    return L


def _prior_apply(op, handle, x):
    """A x on the GPU: Q x minus the low-rank term B Sigma^{-1} B^T x."""
    y = handle.apply_operator(x)
    m = getattr(op, "m_lowrank", 0)
    if m:
        B = op.get_B()
        y = y - B @ ((B.T @ x) / op.get_Sigma())
    return y


def posterior_mean(op, xbar, y, params=None, rtol=1e-12, atol=1e-14, maxiter=200):
    """Posterior mean x_post (LinearOperator::mean, linear_operator.hh:131-147)."""
    xbar = np.asarray(xbar, dtype=np.float64)
    m = getattr(op, "m_lowrank", 0)
    if m == 0:
        return xbar.copy()
    pre = MultigridPreconditioner(op, params or MultigridParameters(nlevel=_nlevel(op)), precision="fp64")
    rhs = _prior_apply(op, pre, xbar) + op.get_B() @ (np.asarray(y, dtype=np.float64) / op.get_Sigma())
    solver = LoopSolver(op, pre, rtol=rtol, atol=atol, maxiter=maxiter)
    return np.asarray(solver.apply(rhs)).reshape(-1)


def _dense(op, b_obs):
    idx, w = b_obs
    b = np.zeros(op.lattice.Nvertex)
    np.add.at(b, np.asarray(idx), np.asarray(w, dtype=np.float64))
    return b


def observed_moments(op, xbar, y, observations, params=None, rtol=1e-12, atol=1e-14, maxiter=200):
    """Means (P,) and covariance (P, P) of z_p = b_p^T x under the posterior for
    observation vectors b_p given as (vertex indices, weights)."""
    x_post = posterior_mean(op, xbar, y, params, rtol, atol, maxiter)
    Bo = np.stack([_dense(op, b) for b in observations])
    W, _ = _solve(op, Bo, params, rtol, atol, maxiter)
    cov = Bo @ W.T
    return Bo @ x_post, 0.5 * (cov + cov.T)


def observed_mean_and_variance(op, xbar, y, b_obs, params=None, rtol=1e-12, atol=1e-14, maxiter=200):
    """(mean, variance) of z = b^T x
    (LinearOperator::observed_mean_and_variance, linear_operator.hh:165-189)."""
    mean, cov = observed_moments(op, xbar, y, [b_obs], params, rtol, atol, maxiter)
    return float(mean[0]), float(cov[0, 0])


def _rhs_for_mean(op, sampler, x_post):
    """f = Q x_post, as the harnesses form it (sampling_time.cc:27)."""
    return np.asarray(sampler.apply_operator(np.tile(x_post, (sampler.nchains, 1))))[0]


def measure_sampling_time(sampler, nwarmup, nsamples, sample_vector, y, label, filename=None, xbar=None,
                          exact=None, verbose=True):
    """Warm-up, timed sampling loop and observable time series
    (measure_sampling_time, src/measurements/sampling_time.cc:4-71).

    The K chains of `sampler` run together from x = 0; the time per sample is
    the wall time of one batched cycle divided by K. The file holds one line
    per sample with one column per chain (one column for K = 1, as in the
    reference). Returns (ms per chain-sample, series[nsamples, K],
    (mean_exact, variance_exact))."""
    op = sampler.get_linear_operator()
    N = op.lattice.Nvertex
    xbar = np.zeros(N) if xbar is None else np.asarray(xbar, dtype=np.float64)
    if exact is None:
        x_post = posterior_mean(op, xbar, y)
        mean_exact, variance_exact = observed_mean_and_variance(op, xbar, y, sample_vector)
    else:
        x_post, mean_exact, variance_exact = exact
    f = _rhs_for_mean(op, sampler, x_post)
    sampler.fix_rhs(f)
    sampler.set_state(np.zeros((sampler.nchains, N)))
    sampler.set_observables([sample_vector])
    sampler.run(nwarmup, record=False)
    t0 = time.perf_counter()
    obs = sampler.run(nsamples)
    t = time.perf_counter() - t0
    data = obs[:, :, 0]
    ms = 1e3 * t / (nsamples * sampler.nchains)
    if filename:
        np.savetxt(filename, data, fmt="%.17g")
    x_avg = data.mean()
    variance = (data * data).mean() - x_avg * x_avg
    x_error = np.sqrt(variance / data.size)
    if verbose:
        print("  %12s time per sample = %12.4f ms" % (label, ms))
        print("  %12s mean     = %12.4e +/- %12.4e [ignoring IACT]" % (label, x_avg, x_error))
        print("  %12s mean     = %12.4e" % ("exact", mean_exact))
        print("  %12s variance = %12.4e" % (label, variance))
        print("  %12s variance = %12.4e\n" % ("exact", variance_exact))
    return ms, data, (mean_exact, variance_exact)


def _batched_series(sampler, f, sample_vector, nsamples, nsteps):
    """z at steps 0..nsteps of ceil(nsamples / K) batches of K chains started
    at x = 0; returns an array (nsamples, nsteps + 1)."""
    N = sampler.get_linear_operator().lattice.Nvertex
    K = sampler.nchains
    sampler.fix_rhs(f)
    sampler.set_observables([sample_vector])
    rows = []
    done = 0
    while done < nsamples:
        sampler.set_state(np.zeros((K, N)))
        obs = sampler.run(nsteps)[:, :, 0] if nsteps else np.zeros((0, K))
        z = np.concatenate([np.zeros((1, K)), obs], axis=0).T  # (K, nsteps + 1); z_0 = b^T 0
        take = min(K, nsamples - done)
        rows.append(z[:take])
        done += take
    return np.concatenate(rows, axis=0)


def convergence_table(z, mean_exact, variance_exact):
    """q_k = |E[z^k] - E[z]| and |Var[z^k] - Var[z]| with their statistical
    errors (convergence.cc:92-113) from z[sample, step]."""
    n = z.shape[0]
    m1, m2, m3, m4 = [(z ** p).mean(axis=0) for p in (1, 2, 3, 4)]
    diff_mean = np.abs(m1 - mean_exact)
    diff_var = np.abs(m2 - m1 * m1 - variance_exact)
    sigma_sq = n / (n - 1.0) * (m2 - m1 * m1)
    mu4 = m4 - 4 * m1 * m3 + 6 * m1 ** 2 * m2 - 3 * m1 ** 4
    err_mean = np.sqrt(sigma_sq / n)
    err_var = np.sqrt((mu4 - (n - 3.0) / (n - 1.0) * sigma_sq * sigma_sq) / n)
    return diff_mean, err_mean, diff_var, err_var


def _convergence_text(diff_mean, err_mean, diff_var, err_var):
    lines = []
    for q in range(2):
        if q == 0:
            lines.append("**** q_k = |E[z^k] - E[z]| **** ")
            label, d, e = "mean", diff_mean, err_mean
        else:
            lines.append("**** q_k = |Var[z^k] - Var[z]| **** ")
            label, d, e = "variance", diff_var, err_var
        lines.append("  %12s   %3s : %12s %35s %35s" % ("", "k", "q_k", "q_k/q_0", "q_k/q_{k-1}"))
        for j in range(len(d)):
            s = "  %12s   %3d : %12.8f +/- %12.8f       %12.8f +/- %12.8f      " % (
                label, j, d[j], e[j], d[j] / d[0], e[j] / d[0])
            if j > 0:
                rel = d[j] / d[j - 1] * np.sqrt((e[j] / d[j]) ** 2 + (e[j - 1] / d[j - 1]) ** 2)
                s += " %12.8f +/- %12.8f " % (d[j] / d[j - 1], rel)
            else:
                s += " %12s" % "---"
            lines.append(s)
        lines.append("")
    return "\n".join(lines) + "\n"


def measure_convergence(sampler, nsamples, nsteps, sample_vector, y, filename=None, xbar=None, exact=None):
    """Convergence of the first two moments of z during warm-up from x = 0
    (measure_convergence, src/measurements/convergence.cc:4-161), over
    `nsamples` independent chains (batches of the sampler's K chains).
    Returns (diff_mean, err_mean, diff_var, err_var) per step 0..nsteps."""
    op = sampler.get_linear_operator()
    N = op.lattice.Nvertex
    xbar = np.zeros(N) if xbar is None else np.asarray(xbar, dtype=np.float64)
    if exact is None:
        x_post = posterior_mean(op, xbar, y)
        mean_exact, variance_exact = observed_mean_and_variance(op, xbar, y, sample_vector)
    else:
        x_post, mean_exact, variance_exact = exact
    f = _rhs_for_mean(op, sampler, x_post)
    z = _batched_series(sampler, f, sample_vector, nsamples, nsteps)
    table = convergence_table(z, mean_exact, variance_exact)
    if filename:
        with open(filename, "w") as fh:
            fh.write(_convergence_text(*table))
    return table


def measure_mean_squared_error(sampler, nsamples, nsteps, sample_vector, y, filename=None, xbar=None, exact=None):
    """Mean squared error of the running chain average of z against the exact
    mean after j steps, j = 0..nsteps-1, over `nsamples` independent chains
    (measure_mean_squared_error, src/measurements/mean_squared_error.cc:4-64).
    Returns (mse[j], error[j]) with the reference's error column."""
    op = sampler.get_linear_operator()
    N = op.lattice.Nvertex
    xbar = np.zeros(N) if xbar is None else np.asarray(xbar, dtype=np.float64)
    if exact is None:
        x_post = posterior_mean(op, xbar, y)
        mean_exact, _ = observed_mean_and_variance(op, xbar, y, sample_vector)
    else:
        x_post, mean_exact, _ = exact
    f = _rhs_for_mean(op, sampler, x_post)
    z = _batched_series(sampler, f, sample_vector, nsamples, max(nsteps - 1, 0))[:, :nsteps]
    running = np.cumsum(z, axis=1) / np.arange(1, nsteps + 1)
    sq = (running - mean_exact) ** 2
    mse = sq.mean(axis=0)
    mse_error = 1.0 / (nsamples - 1.0) * ((sq * sq).mean(axis=0) - mse * mse)
    if filename:
        with open(filename, "w") as fh:
            for j in range(nsteps):
                fh.write(" %5d : %12.6e +/- %12.6e\n" % (j, mse[j], mse_error[j]))
    return mse, mse_error
