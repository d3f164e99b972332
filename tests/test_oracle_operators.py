"""Pins the CPU oracle's lattice, operator, measurement and intergrid pieces.

Each check is against something the oracle's code does not itself compute:
the reference's manufactured-solution tests, multilinear interpolation built
from vertex coordinates, dense brute-force Galerkin products, and closed-form
properties of the ball-average measurement vectors.
"""
import itertools

import numpy as np
import pytest


def _f(z):  # test_linear_operator.hh:43-47
    return 100 * z * z * (1 - z) * np.exp(-6 * z)


def _d2f(z):  # test_linear_operator.hh:54-58
    return 100 * (2 + z * (-30 + z * (72 - 36 * z))) * np.exp(-6 * z)


def _vertex_coords(n):
    """Coordinates of interior vertices in reference order (x fastest)."""
    axes = [np.arange(1, nd) / nd for nd in n]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    return np.stack([g.reshape(-1) for g in grids], axis=1)


@pytest.mark.parametrize("dim,n,tol", [(2, (512, 512), 2e-4), (3, (64, 64, 64), 7e-3)])
def test_fd_manufactured_solution(oracle, dim, n, tol):
    """test_linear_operator.hh:212-246 (periodic Lambda in [1.3, 2.3])."""
    op = oracle.Operator.fd_periodic(dim, n, 1.3, 2.3)
    X = _vertex_coords(n)
    L1, L2 = 0.5 * (2.3 + 1.3), 0.5 * (2.3 - 1.3)
    lam = L1 + L2 * np.prod(np.cos(np.pi * X), axis=1)
    kappa_sq = 1.0 / lam**2
    u = np.prod(_f(X), axis=1)
    rhs = kappa_sq * u
    for j in range(dim):
        dd = np.ones(len(X))
        for d in range(dim):
            dd *= _d2f(X[:, d]) if j == d else _f(X[:, d])
        rhs -= dd
    rhs *= np.prod([1.0 / v for v in n])
    Au = op.apply(u)
    assert np.linalg.norm(Au - rhs) / np.linalg.norm(Au) < tol


def test_fd_symmetric_and_scaled(oracle):
    op = oracle.Operator.fd(3, (8, 8, 8), 10.0)
    A = op.csr(0).toarray()
    assert np.allclose(A, A.T, rtol=0, atol=0)
    h = 1 / 8
    # operator annihilates nothing: constant field sees kappa^2 h^3 in the bulk
    c = np.ones(op.N)
    y = op.apply(c)
    X = _vertex_coords((8, 8, 8))
    interior = np.all((X > 1.5 * h) & (X < 1 - 1.5 * h), axis=1)
    assert np.allclose(y[interior], 10.0 * h**3, rtol=1e-12)


def _interp_matrix(n):
    """Dense multilinear interpolation from coarse (n/2) to fine (n) interior vertices."""
    nc = tuple(v // 2 for v in n)
    Xf = _vertex_coords(n)
    Xc = _vertex_coords(nc)
    P = np.ones((len(Xf), len(Xc)))
    for d in range(len(n)):
        H = 1.0 / nc[d]
        P *= np.maximum(0.0, 1.0 - np.abs(Xf[:, d:d + 1] - Xc[None, :, d]) / H)
    return P


@pytest.mark.parametrize("dim,n", [(1, (16,)), (2, (8, 16)), (3, (8, 8, 8))])
def test_prolongation_is_multilinear_interpolation(oracle, dim, n):
    """test_intergrid.hh:87-146: prolongate_add equals interpolation."""
    op = oracle.Operator.fd(dim, n, 1.0)
    P = _interp_matrix(n)
    rng = np.random.default_rng(7)
    xc = rng.standard_normal(P.shape[1])
    x0 = rng.standard_normal(P.shape[0])
    out = op.prolongate_add(0.7, xc, x0)
    assert np.allclose(out, x0 + 0.7 * P @ xc, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim,n", [(2, (16, 16)), (3, (8, 8, 8))])
def test_restriction_is_adjoint(oracle, dim, n):
    """test_intergrid.hh:155-171: <P u, r> = <u, R r>."""
    op = oracle.Operator.fd(dim, n, 1.0)
    rng = np.random.default_rng(11)
    u = rng.uniform(size=op.coarse_ndof())
    r = rng.uniform(size=op.N)
    Pu = op.prolongate_add(1.0, u, np.zeros(op.N))
    assert abs(Pu @ r - u @ op.restrict(r)) < 1e-12


@pytest.mark.parametrize("dim,n", [(2, (16, 16)), (3, (8, 8, 8))])
def test_galerkin_coarsening_bruteforce(oracle, dim, n):
    """A_c = P^T A P and B_c = P^T B (linear_operator.cc:10-23), dense brute force."""
    op = oracle.Operator.fd(dim, n, 10.0)
    rng = np.random.default_rng(3)
    centres = rng.uniform(0.2, 0.8, size=(3, dim))
    op.add_measurements(centres, 0.15, np.array([1e-4, 2e-4, 3e-4]))
    P = _interp_matrix(n)
    A = op.csr(0).toarray()
    B = op.csr(1).toarray()
    co = op.coarsen()
    assert np.allclose(co.csr(0).toarray(), P.T @ A @ P, rtol=0, atol=1e-14)
    assert np.allclose(co.csr(1).toarray(), P.T @ B, rtol=0, atol=1e-14)
    assert np.array_equal(co.sigma(), op.sigma())
    # constant-coefficient FD coarsens to a translation-invariant box stencil
    Ac = co.csr(0).toarray()
    nc = tuple(v // 2 for v in n)
    Xc = _vertex_coords(nc)
    deep = np.where(np.all((Xc > 1.5 / np.array(nc)) & (Xc < 1 - 1.5 / np.array(nc)), axis=1))[0]
    rows = Ac[deep]
    assert np.allclose(np.sort(rows, axis=1), np.sort(rows[0])[None, :], rtol=1e-12)
    assert np.count_nonzero(rows[0]) == 3**dim


def test_measurement_vector_ball_average(oracle):
    """Top-hat ball average (measured_operator.cc:69-170): sum of weights -> 1,
    centroid -> ball centre, support inside the ball's cell neighbourhood."""
    x0 = np.array([0.43, 0.51, 0.47])
    errs = []
    for n in (32, 64):
        op = oracle.Operator.fd(3, (n, n, n), 1.0)
        b = op.measurement_vector(x0, 0.1)
        X = _vertex_coords((n, n, n))
        s = b.sum()
        errs.append(abs(s - 1))
        centroid = (b[:, None] * X).sum(axis=0) / s
        assert np.allclose(centroid, x0, atol=2.0 / n)
        dist = np.linalg.norm(X[b != 0] - x0, axis=1)
        assert dist.max() < 0.1 + np.sqrt(3) / n
    assert errs[-1] < 0.05 and errs[-1] < errs[0]


def test_measurement_vector_nearest_vertex(oracle):
    """radius < 1e-12 selects the nearest vertex (measured_operator.cc:74-90)."""
    op = oracle.Operator.fd(2, (16, 16), 1.0)
    b = op.measurement_vector(np.array([0.30, 0.72]), 0.0)
    X = _vertex_coords((16, 16))
    k = np.argmin(np.linalg.norm(X - [0.30, 0.72], axis=1))
    e = np.zeros(op.N)
    e[k] = 1.0
    assert np.array_equal(b, e)


def test_posterior_apply_is_dense_precision(oracle):
    op = oracle.Operator.fd(2, (16, 16), 10.0)
    op.add_measurements(np.array([[0.3, 0.3], [0.6, 0.7]]), 0.1, np.array([1e-3, 2e-3]), variance_scaling=0.5)
    A = op.csr(0).toarray()
    B = op.csr(1).toarray()
    Q = A + B @ np.diag(1 / op.sigma()) @ B.T
    assert np.allclose(op.sigma(), 0.5 * np.array([1e-3, 2e-3]))
    assert np.allclose(op.dense_precision(), Q, rtol=1e-13, atol=1e-16)
    x = np.random.default_rng(0).standard_normal(op.N)
    assert np.allclose(op.apply(x), Q @ x, rtol=1e-12, atol=1e-14)


def test_lattice_errors(oracle):
    op = oracle.Operator.fd(2, (6, 6), 1.0)
    op3 = op.coarsen()  # 6 -> 3
    with pytest.raises(RuntimeError):
        op3.coarsen()  # odd extent
    with pytest.raises(RuntimeError):
        oracle.Operator.fd(2, (4, 4), 1.0).coarsen().coarsen()  # no interior vertices
