"""Shared set-up for the GPU parity tests: one problem built twice, once for
the HIP library (multigridmc_amd) and once for the CPU oracle."""
import numpy as np

import multigridmc_amd as mg
from oracle import pyoracle as po


def make_problem(dim, n, nmeas, kappa_sq=10.0, radius=0.15, var=1e-3, seed=17):
    lat = mg.Lattice(*n)
    prior = mg.ShiftedLaplaceFDOperator(lat, kappa_sq)
    ora = po.Operator.fd(dim, n, kappa_sq)
    if nmeas:
        rng = np.random.default_rng(seed)
        centres = rng.uniform(0.2, 0.8, size=(nmeas, dim))
        variance = var * (1.0 + rng.uniform(size=nmeas))
        op = mg.MeasuredOperator(prior, centres, radius, variance)
        ora.add_measurements(centres, radius, variance)
    else:
        op = prior
    return op, ora


def params(nlevel, smoother="SSOR", omega=1.0, cycle=1, coarse="Cholesky"):
    return mg.MultigridParameters(nlevel=nlevel, smoother=smoother, omega=omega, cycle=cycle, coarse_solver=coarse)


def oracle_mg(ora, p, ordering=po.MULTICOLOUR):
    return po.Multigrid(ora, p.nlevel, smoother=0 if p.smoother == "SOR" else 1, npresmooth=p.npresmooth,
                        npostsmooth=p.npostsmooth, ncoarsesmooth=p.ncoarsesmooth, omega=p.omega, cycle=p.cycle,
                        coarse_scaling=p.coarse_scaling, coarse_solver=0 if p.coarse_solver == "Cholesky" else 1,
                        ordering=ordering)


# ---- Philox4x32-10 and Box-Muller in NumPy (independent of the C++ header) ----
M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(v, dtype=np.uint64) & MASK for v in (c0, c1, c2, c3)]
    k0 = np.uint64(k0)
    k1 = np.uint64(k1)
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
            k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return c


def u24(b):
    return (b >> np.uint64(8)).astype(np.float64) / 16777216.0 + 1.0 / 33554432.0


def normals_reference(lat, nchains, tag, sample, seed):
    """Normal at every fine vertex for every chain, following the documented
    slot layout: ctr0 = ((row * px/8 + i/8) << 1) | (i & 1), lane (i & 7) >> 1."""
    n = lat.n
    dim = lat.dim
    px = ((n[0] + 7) // 8) * 8
    py = n[1] + 1
    N = lat.Nvertex
    idx = lat.vertex_euclidean(np.arange(N))
    i, j = idx[:, 0], idx[:, 1]
    k = idx[:, 2] if dim == 3 else np.zeros_like(i)
    ctr0 = (((k * py + j) * (px // 8) + (i >> 3)) << 1) | (i & 1)
    lane = (i & 7) >> 1
    out = np.zeros((nchains, N))
    for c in range(nchains):
        r = philox(ctr0, c, tag, sample, seed & MASK, seed >> 32)
        ra = np.sqrt(-2.0 * np.log(u24(r[0])))
        rb = np.sqrt(-2.0 * np.log(u24(r[2])))
        ua, ub = u24(r[1]), u24(r[3])
        z = np.stack([ra * np.cos(2 * np.pi * ua), ra * np.sin(2 * np.pi * ua), rb * np.cos(2 * np.pi * ub),
                      rb * np.sin(2 * np.pi * ub)], axis=1)
        out[c] = z[np.arange(N), lane]
    return out


# ---- statistical comparison of sampled probe moments with exact ones ----
def moment_zscores(obs, mean_exact, cov_exact):
    """obs[nsamples, nchains, nprobe] -> (z of the probe means, z of the
    covariance entries on and above the diagonal): each deviation from the
    exact value in units of its Monte Carlo standard error. A mean's error is
    sqrt(tau_p C_pp / n); a covariance entry, estimated about the exact mean,
    has error sqrt(tau (C_pp C_qq + C_pq^2) / n) for Gaussian samples, with
    tau the larger of its probes' integrated autocorrelation times (an upper
    bound for the product series) and n = nsamples * nchains."""
    n, K, P = obs.shape
    taus = np.array([max(1.0, mg.tau_int_chains(obs[:, :, p], mean=mean_exact[p])[0]) for p in range(P)])
    z = obs.reshape(-1, P)
    ns = z.shape[0]
    dg = np.diag(cov_exact)
    zm = (z.mean(axis=0) - mean_exact) / np.sqrt(dg * taus / ns)
    d = z - mean_exact
    C = d.T @ d / ns
    se = np.sqrt((np.outer(dg, dg) + cov_exact ** 2) * np.maximum.outer(taus, taus) / ns)
    iu = np.triu_indices(P)
    return zm, ((C - cov_exact) / se)[iu]


def familywise_z(nquantities, alpha=0.01):
    """Two-sided bound that nquantities standard normal deviates all stay
    within with probability at least 1 - alpha (Bonferroni)."""
    from scipy.stats import norm
    return float(norm.isf(alpha / (2.0 * nquantities)))
