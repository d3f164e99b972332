"""Integrated autocorrelation time estimators for the MGMC metric.

* tau_int_window: the reference estimator Statistics::tau_int
  (src/auxilliary/statistics.cc:4-79): tau = 1 + 2 sum_{k=1}^{K-1} (1 - k/K) C(k)/C(0)
  with C(k) = S_k - mean^2 and S_k the average of q_n q_{n-k} over the n - k
  available pairs, for a fixed window K.
* tau_int_sokal: Sokal's automatic windowing (window W = smallest W with
  W >= c * tau(W), c = 5) on the FFT autocovariance.
"""
import numpy as np


def tau_int_window(q, window):
    """The window is clamped to the length of the series (a series of at
    most one sample gives 1)."""
    q = np.asarray(q, dtype=np.float64)
    n = len(q)
    window = min(int(window), n)
    if window <= 1:
        return 1.0
    avg = q.mean()
    S = np.array([np.dot(q[k:], q[: n - k]) / (n - k) for k in range(window)])
    C = S - avg * avg
    k = np.arange(1, window)
    return 1.0 + 2.0 * np.sum((1.0 - k / window) * C[1:] / C[0])


def _autocov(q):
    q = np.asarray(q, dtype=np.float64)
    n = len(q)
    d = q - q.mean()
    nfft = 1 << (2 * n - 1).bit_length()
    F = np.fft.rfft(d, nfft)
    ac = np.fft.irfft(F * np.conj(F), nfft)[:n] / n
    return ac


def tau_int_sokal(q, c=5.0):
    """Returns (tau, window)."""
    ac = _autocov(q)
    if ac[0] <= 0:
        return 1.0, 0
    rho = ac / ac[0]
    tau = 1.0
    for W in range(1, len(rho)):
        tau += 2.0 * rho[W]
        if W >= c * tau:
            return tau, W
    return tau, len(rho) - 1


def tau_int_chains(obs, c=5.0, mean=None):
    """obs[nsamples, nchains] -> (tau, window) from the chain-averaged
    autocovariance (chains are independent and identically distributed).

    Deviations are taken from one mean for all chains: the exact mean when
    `mean` is given, else the mean pooled over every sample of every chain.
    Subtracting each chain's own mean instead removes that chain's
    low-frequency part and reads low on short chains (about 0.7 for a true
    1.15 at 20 steps per chain); the pooled mean keeps that bias of order
    tau / (nsamples * nchains)."""
    obs = np.asarray(obs, dtype=np.float64)
    n, K = obs.shape
    d = obs - (obs.mean() if mean is None else float(mean))
    nfft = 1 << (2 * n - 1).bit_length()
    F = np.fft.rfft(d, nfft, axis=0)
    ac = np.fft.irfft(F * np.conj(F), nfft, axis=0)[:n].sum(axis=1) / (n * K)
    rho = ac / ac[0]
    tau = 1.0
    for W in range(1, n):
        tau += 2.0 * rho[W]
        if W >= c * tau:
            return tau, W
    return tau, n - 1
