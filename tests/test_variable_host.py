"""Variable coefficients on the host (no device work): the per-vertex
stencils the library builds for -div(a grad) + kappa^2(x) and their
Galerkin coarsenings (mgmc_level_stencils), against

* the same operator assembled in Python as a CSR matrix (bit for bit on the
  fine level), and the oracle's restatement of the reference's periodic
  correlation-length model (src/linear_operator/correlationlength_model.hh:83-105,
  shiftedlaplace_fd_operator.cc:33-56);
* the oracle's Galerkin coarsening R A R^T (sparse triple product) and a
  dense P^T A P on an 8^3 lattice;
* the constant-coefficient reduction: a == c everywhere gives the stencils
  of h^d (-c Delta + kappa^2), the constant path's."""
import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import _lib
from multigridmc_amd.operators import block_diffusion, periodic_kappa_sq
from multigridmc_amd.sampler import _problem


def level_stencils(op, nlevel, level):
    prob, keep = _problem(op)
    N = mg.Lattice(*[v >> level for v in op.lattice.n]).Nvertex
    out = np.zeros((N, 27))
    _lib.check(_lib.lib().mgmc_level_stencils(prob, nlevel, level, out.ctypes.data))
    return out


def csr_stencils(A, lattice):
    """Rows of a CSR matrix as (N, 27) stencils in the library's index order."""
    A = A.tocsr()
    N = lattice.Nvertex
    idx = lattice.vertex_euclidean(np.arange(N))
    out = np.zeros((N, 27))
    rows = np.repeat(np.arange(N), np.diff(A.indptr))
    o = lattice.vertex_euclidean(A.indices) - idx[rows]
    o3 = np.zeros((len(rows), 3), dtype=np.int64)
    o3[:, :lattice.dim] = o
    e = (o3[:, 2] + 1) * 9 + (o3[:, 1] + 1) * 3 + (o3[:, 0] + 1)
    np.add.at(out, (rows, e), A.data)
    return out


CASES = [(2, (16, 16), (4, 4)), (2, (32, 16), (8, 4)), (3, (8, 8, 8), (2, 2, 2)), (3, (16, 16, 16), (8, 8, 8))]


@pytest.mark.parametrize("dim,n,blocks", CASES)
def test_fine_stencils_match_python_assembly(dim, n, blocks):
    lat = mg.Lattice(*n)
    a = block_diffusion(lat, blocks, 1.0, 10.0, 5)
    kap = 10.0 + np.random.default_rng(1).uniform(size=lat.Nvertex)
    op = mg.ShiftedLaplaceFDOperator(lat, kap, diffusion=a)
    assert np.array_equal(level_stencils(op, 2, 0), csr_stencils(op.get_sparse(), lat))


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (16, 16, 16))])
def test_periodic_kappa_matches_reference_model(dim, n, oracle):
    lat = mg.Lattice(*n)
    op = mg.ShiftedLaplaceFDOperator(lat, periodic_kappa_sq(lat, 1.3, 2.3))
    ref = csr_stencils(oracle.Operator.fd_periodic(dim, n, 1.3, 2.3).csr(0), lat)
    got = level_stencils(op, 2, 0)
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()


@pytest.mark.parametrize("dim,n,blocks", CASES)
def test_galerkin_stencils_match_oracle(dim, n, blocks, oracle):
    lat = mg.Lattice(*n)
    a = block_diffusion(lat, blocks, 1.0, 10.0, 7)
    op = mg.ShiftedLaplaceFDOperator(lat, 10.0, diffusion=a)
    A = op.get_sparse().tocoo()
    ora = oracle.Operator.custom(dim, n, A.row, A.col, A.data)
    nlevel = 3 if min(n) >= 16 else 2
    for level in range(1, nlevel):
        ora = ora.coarsen()
        latc = mg.Lattice(*[v >> level for v in n])
        ref = csr_stencils(ora.csr(0), latc)
        got = level_stencils(op, nlevel, level)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), level


def test_galerkin_stencils_dense_brute_force():
    """8^3 with a random cell-wise coefficient: P^T A P formed densely."""
    lat = mg.Lattice(8, 8, 8)
    a = np.random.default_rng(3).uniform(1.0, 10.0, size=512)
    op = mg.ShiftedLaplaceFDOperator(lat, 10.0, diffusion=a)
    A = op.get_sparse().toarray()
    latc = mg.Lattice(4, 4, 4)
    Xf = lat.vertex_euclidean(np.arange(lat.Nvertex))
    Xc = latc.vertex_euclidean(np.arange(latc.Nvertex))
    P = np.ones((lat.Nvertex, latc.Nvertex))
    for d in range(3):
        dist = np.abs(Xf[:, None, d] - 2 * Xc[None, :, d])
        P *= np.where(dist == 0, 1.0, np.where(dist == 1, 0.5, 0.0))
    ref = csr_stencils(__import__("scipy.sparse", fromlist=["csr_matrix"]).csr_matrix(P.T @ A @ P), latc)
    got = level_stencils(op, 2, 1)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (16, 16, 16))])
def test_constant_diffusion_field_reduces_to_constant_operator(dim, n):
    lat = mg.Lattice(*n)
    c = 3.7
    field = mg.ShiftedLaplaceFDOperator(lat, 10.0, diffusion=np.full(int(np.prod(n)), c))
    const = mg.ShiftedLaplaceFDOperator(lat, 10.0, diffusion=c)
    for level in range(3):
        got, ref = level_stencils(field, 3, level), level_stencils(const, 3, level)
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), level
