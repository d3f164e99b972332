"""Pins the oracle's CLCG generator and the reference IACT estimator."""
import math

import numpy as np


def test_clcg_recurrence_first_draws(oracle):
    """L'Ecuyer's combined LCG from seeds (1, 1) (parallel_random.cc:35-70):
    x_0 = 1 - 1 -> m1 - 1; next states 40014 and 40692."""
    m1, a1, m2, a2 = 2147483563, 40014, 2147483399, 40692
    y1, y2, want = 1, 1, []
    for _ in range(5):
        x = y1 - y2
        if x < 1:
            x += m1 - 1
        want.append(x)
        y1, y2 = (a1 * y1) % m1, (a2 * y2) % m2
    got = oracle.CLCG().ints(5)
    assert list(got) == want
    assert got[0] == m1 - 1


def test_clcg_leapfrog_consistent(oracle):
    """test_random.hh:27-51: four leapfrogged streams interleave to the serial one."""
    serial = oracle.CLCG().ints(64)
    par = np.zeros(64, np.int32)
    for t in range(4):
        par[t::4] = oracle.CLCG(4, t).ints(16)
    assert np.array_equal(serial, par)


def test_clcg_uniform_cdf(oracle):
    """test_random.hh:54-79: 2^24 uniforms, 32 bins, CDF error < 2e-4."""
    u = oracle.CLCG().uniform(1 << 24)
    h, _ = np.histogram(u, bins=32, range=(0, 1))
    cdf = np.cumsum(h) / len(u)
    assert np.abs(cdf - np.arange(1, 33) / 32).max() < 2e-4


def test_clcg_normal_cdf(oracle):
    """test_random.hh:82-109: 2^24 polar-method normals, 64 bins on [-10, 10]."""
    z = oracle.CLCG().normal(1 << 24)
    h, edges = np.histogram(z, bins=64, range=(-10, 10))
    cdf = np.cumsum(h) / len(z)
    true = np.array([0.5 * (1 + math.erf(x / math.sqrt(2))) for x in edges[1:]])
    assert np.abs(cdf - true).max() < 2e-4


def test_tau_int_ar_process(oracle):
    """test_statistics.hh:145-166: Q_n = A Q_{n-1} + xi + v with a closed-form
    autocovariance; window 20, direction v = (0.8, 0) as the reference sets it."""
    theta = 1.3
    rot = np.array([[math.cos(theta), math.sin(theta)], [-math.sin(theta), math.cos(theta)]])
    A = rot @ np.diag([0.6, 0.4]) @ rot.T
    shift = np.array([1.4, 0.6])
    rng = np.random.default_rng(1241517)
    n, nwarm = 1_000_000, 1000
    xi = rng.standard_normal((n + nwarm, 2))
    Q = np.zeros(2)
    out = np.empty((n, 2))
    for j in range(n + nwarm):
        Q = A @ Q + xi[j] + shift
        if j >= nwarm:
            out[j - nwarm] = Q
    v = np.array([0.8, 0.0])
    window = 20
    w = np.linalg.solve(np.eye(2) - A @ A, v)
    wk = w.copy()
    tau_exact = 1.0
    for k in range(1, window):
        wk = A @ wk
        tau_exact += 2.0 * (1.0 - k / window) * (v @ wk) / (v @ w)
    tau = oracle.tau_int_reference(out @ v, window)
    assert abs(tau - tau_exact) < 2e-2
