"""CPU-side checks of the product package: the C ABI library loads and exports
every declared entry point, host-side operator setup equals the oracle's, and
argument errors come back as status codes (no GPU work is issued)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import _lib
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mgmc_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(\w+)\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("dim,n,m,radius", [(2, (32, 32), 5, 0.1), (3, (16, 32, 16), 4, 0.12), (2, (16, 16), 3, 0.0)])
def test_measurement_matrix_matches_oracle(dim, n, m, radius):
    lat = mg.Lattice(*n)
    rng = np.random.default_rng(11)
    centres = rng.uniform(0.1, 0.9, size=(m, dim))
    var = rng.uniform(1e-4, 2e-4, size=m)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 7.0), centres, radius, var, variance_scaling=0.5)
    ora = po.Operator.fd(dim, n, 7.0)
    ora.add_measurements(centres, radius, var, variance_scaling=0.5)
    assert abs(op.get_B() - ora.csr(1)).max() < 1e-15
    assert abs(op.get_sparse() - ora.csr(0)).max() == 0.0
    assert np.array_equal(op.get_Sigma(), ora.sigma())
    x = rng.standard_normal(lat.Nvertex)
    assert np.allclose(op.apply(x), ora.apply(x), rtol=1e-13, atol=1e-15)


def _create(n=(16, 16, 16), nlevel=3, **kw):
    lat = mg.Lattice(*n)
    op = mg.ShiftedLaplaceFDOperator(lat, 10.0)
    p = mg.MultigridParameters(nlevel=nlevel, **kw)
    return mg.MultigridMCSampler(op, p)


@pytest.mark.parametrize("kw,code", [
    (dict(n=(18, 18, 18), nlevel=3), 1),          # odd extent after one coarsening
    (dict(n=(8, 8, 8), nlevel=4), 1),             # no interior vertices
    (dict(omega=2.5), 1),                         # omega outside (0, 2)
    (dict(smoother="Jacobi"), 1),                 # invalid smoother
    (dict(n=(64, 64, 64), nlevel=2), 2),          # coarsest level too large for dense Cholesky
])
def test_invalid_arguments_return_status(kw, code):
    with pytest.raises(mg.MGMCError) as e:
        _create(**kw)
    assert e.value.code == code


def test_non_stencil_operator_is_rejected():
    lat = mg.Lattice(8, 8)
    N = lat.Nvertex
    rows = np.arange(N)
    A = np.zeros((N, N))
    A[rows, rows] = 4.0
    A[0, N - 1] = A[N - 1, 0] = -1.0  # couples far-apart vertices
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    prob = _lib.Problem()
    prob.dim = 2
    prob.n[0], prob.n[1], prob.n[2] = 8, 8, 1
    rp, ci, v = _lib.i64(A.indptr), _lib.i64(A.indices), _lib.f64(A.data)
    prob.A_rowptr, prob.A_col, prob.A_val = _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(v)
    par = _lib.Params(nlevel=2, smoother=1, npresmooth=1, npostsmooth=1, ncoarsesmooth=1, omega=1.0, cycle=1,
                      coarse_scaling=1.0, coarse_solver=0, precision=0, nchains=1, seed=1, stream=None)
    h = C.c_void_p()
    code = _lib.lib().mgmc_create(C.byref(prob), C.byref(par), C.byref(h))
    assert code == 2
    assert b"neighbours" in _lib.lib().mgmc_last_error()


def test_stats_estimators_on_ar1():
    """tau_int of an AR(1) series x_n = a x_{n-1} + e_n is (1 + a)/(1 - a)."""
    a = 0.6
    rng = np.random.default_rng(0)
    n = 400000
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0]
    for i in range(1, n):
        x[i] = a * x[i - 1] + e[i]
    tau_exact = (1 + a) / (1 - a)
    tau, W = mg.tau_int_sokal(x)
    assert abs(tau - tau_exact) < 0.1
    # windowed estimator equals the reference's running-average form
    w = 12
    assert abs(mg.tau_int_window(x[:20000], w) - po.tau_int_reference(x[:20000], w)) < 1e-9
    # a window longer than the series is clamped to its length
    assert mg.tau_int_window(x[:5], 10) == mg.tau_int_window(x[:5], 5)
    assert mg.tau_int_window(x[:1], 10) == 1.0
    chains = np.stack([x[i * 40000:(i + 1) * 40000] for i in range(10)], axis=1)
    tau_c, _ = mg.tau_int_chains(chains)
    assert abs(tau_c - tau_exact) < 0.15


def _ar1_chains(a, nsteps, nchains, rng):
    """Stationary AR(1) chains x_n = a x_{n-1} + sqrt(1 - a^2) e_n (unit variance)."""
    x = np.empty((nsteps, nchains))
    x[0] = rng.standard_normal(nchains)
    s = np.sqrt(1 - a * a)
    for i in range(1, nsteps):
        x[i] = a * x[i - 1] + s * rng.standard_normal(nchains)
    return x


def test_tau_int_chains_unbiased_on_short_chains():
    """The batched estimator pools the mean over all chains, so 64 chains of
    20 steps (a short timed run) give the same tau as 400 steps: AR(1) with
    tau = 1.15, the MGMC value at 128^3. A per-chain mean reads about 0.7."""
    a = 0.15 / 2.15
    tau_exact = (1 + a) / (1 - a)
    rng = np.random.default_rng(3)
    for nsteps, tol in ((20, 0.06), (400, 0.03)):
        t = [mg.tau_int_chains(_ar1_chains(a, nsteps, 64, rng))[0] for _ in range(40)]
        assert abs(np.median(t) - tau_exact) < tol, (nsteps, np.median(t))
    # with the exact mean supplied the estimate is the same to within noise
    x = _ar1_chains(a, 400, 64, rng)
    assert abs(mg.tau_int_chains(x, mean=0.0)[0] - mg.tau_int_chains(x)[0]) < 0.02


def test_setup_helpers_match_host_operators():
    """mgmc_measurement_vector and mgmc_fem_stencil (C++ set-up code in the
    library, no device work) equal the Python operators bit for bit."""
    import ctypes as C
    from multigridmc_amd import _lib
    L = _lib.lib()
    for dim, n, x0, r in [(3, (16, 16, 16), (0.3, 0.55, 0.61), 0.15), (2, (64, 32), (0.41, 0.52), 0.07),
                          (3, (32, 32, 32), (0.2, 0.8, 0.45), 0.05), (3, (16, 16, 16), (0.3, 0.55, 0.61), 0.0)]:
        idx, val = mg.measurement_vector(mg.Lattice(*n), x0, r)
        nn = np.array(list(n) + [1] * (3 - len(n)), dtype=np.int32)
        cap = 4096
        oi, ov, nnz = np.zeros(cap, np.int64), np.zeros(cap), C.c_int64()
        x0a = np.array(x0, dtype=np.float64)
        _lib.check(L.mgmc_measurement_vector(dim, nn.ctypes.data, x0a.ctypes.data, r, cap, oi.ctypes.data,
                                             ov.ctypes.data, C.byref(nnz)))
        assert nnz.value == len(idx) and np.array_equal(oi[:len(idx)], idx) and np.array_equal(ov[:len(idx)], val)
    for dim, n in [(2, (16, 32)), (3, (8, 16, 32))]:
        st = np.zeros(27 if dim == 3 else 9)
        nn = np.array(list(n) + [1] * (3 - len(n)), dtype=np.int32)
        _lib.check(L.mgmc_fem_stencil(dim, nn.ctypes.data, 10.0, st.ctypes.data))
        ref = mg.ShiftedLaplaceFEMOperator(mg.Lattice(*n), 10.0).stencil()
        for s, v in ref.items():
            e = ((s[2] + 1) * 9 if dim == 3 else 0) + (s[1] + 1) * 3 + (s[0] + 1)
            assert abs(st[e] - v) <= 1e-15 * abs(v)
