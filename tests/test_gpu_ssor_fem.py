"""Standalone SSOR Gibbs sampler (src/sampler/ssor_sampler.cc) and the Q1 FEM
prior (src/linear_operator/shiftedlaplace_fem_operator.cc) on the GPU."""
import numpy as np
import pytest

import multigridmc_amd as mg
from gpu_common import oracle_mg, params
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _fem_problem(dim, n, nmeas, seed=3, radius=0.15, var=1e-3):
    lat = mg.Lattice(*n)
    prior = mg.ShiftedLaplaceFEMOperator(lat, 10.0)
    A = prior.get_sparse().tocoo()
    ora = po.Operator.custom(dim, n, A.row, A.col, A.data)
    if nmeas:
        rng = np.random.default_rng(seed)
        centres = rng.uniform(0.2, 0.8, size=(nmeas, dim))
        variance = var * (1.0 + rng.uniform(size=nmeas))
        op = mg.MeasuredOperator(prior, centres, radius, variance)
        ora.add_measurements(centres, radius, variance)
        return op, ora
    return prior, ora


def _stencil_array(st, dim):
    out = np.zeros(27 if dim == 3 else 9)
    for s, v in st.items():
        if dim == 3:
            out[(s[2] + 1) * 9 + (s[1] + 1) * 3 + (s[0] + 1)] = v
        else:
            out[(s[1] + 1) * 3 + (s[0] + 1)] = v
    return out


@pytest.mark.parametrize("dim,n", [(2, (32, 32)), (3, (16, 16, 16)), (3, (16, 32, 16))])
def test_fem_coarsening_reproduces_coarse_fem(dim, n):
    """Galerkin coarsening of the FEM operator is the FEM operator of the
    coarse lattice (test_intergrid.hh:179-207, tolerance 1e-12)."""
    prior = mg.ShiftedLaplaceFEMOperator(mg.Lattice(*n), 10.0)
    s = mg.MultigridMCSampler(prior, params(3))
    lat = mg.Lattice(*n)
    for level in range(3):
        got, _ = s.level_stencil(level)
        want = _stencil_array(mg.ShiftedLaplaceFEMOperator(lat, 10.0).stencil(), dim)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
        lat = lat.get_coarse_lattice()


@pytest.mark.parametrize("dim,n,nlevel", [(2, (32, 32), 3), (3, (16, 16, 16), 2)])
def test_fem_posterior_mg_solve_matches_oracle(dim, n, nlevel):
    """driver_mg path on a FEM-prior posterior: residual history equal to the
    oracle's multicolour twin, solution to 1e-10 of a dense solve."""
    op, ora = _fem_problem(dim, n, 3, radius=0.1, var=1e-4)
    p = params(nlevel)
    b = np.random.default_rng(17).standard_normal(op.lattice.Nvertex)
    # iteration counts at a tolerance well above the rounding floor
    solver = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p), rtol=1e-8, atol=1e-14, maxiter=60)
    solver.apply(b)
    _, hist_ref, it_ref = oracle_mg(ora, p).loop_solve(b, rtol=1e-8, atol=1e-14, maxiter=60)
    assert solver.niter == it_ref
    assert np.allclose(solver.history, hist_ref[: len(solver.history)], rtol=1e-8)
    solver = mg.LoopSolver(op, mg.MultigridPreconditioner(op, p), rtol=1e-12, atol=1e-14, maxiter=60)
    x = solver.apply(b)
    x_dense = ora.dense_solve(b)
    assert np.linalg.norm(x - x_dense) / np.linalg.norm(x_dense) < 1e-10


def _exact(ora, f):
    cov = np.linalg.inv(ora.dense_precision())
    return cov @ f, cov


@pytest.mark.parametrize("kind", ["ssor_fd", "ssor_fem", "mgmc_fem"])
def test_sampler_moments(kind):
    """Stationary mean and variances of K independent chains against the
    exact posterior (dense), within 5 Monte Carlo standard errors."""
    dim, n = 2, (16, 16)
    if kind.endswith("fem"):
        op, ora = _fem_problem(dim, n, 2)
    else:
        from gpu_common import make_problem
        op, ora = make_problem(dim, n, 2)
    N = op.lattice.Nvertex
    f = np.random.default_rng(4).standard_normal(N) * 0.01
    K = 4096
    if kind.startswith("ssor"):
        s = mg.SSORSampler(op, omega=1.0, nsmooth=1, nchains=K, seed=77)
        nwarm = 120
    else:
        s = mg.MultigridMCSampler(op, params(3), nchains=K, seed=77)
        nwarm = 12
    s.fix_rhs(f)
    s.set_state(np.zeros((K, N)))
    s.run(nwarm, record=False)
    x = s.get_state()
    mean, cov = _exact(ora, f)
    var = np.diag(cov)
    assert np.all(np.abs(x.mean(axis=0) - mean) < 5 * np.sqrt(var / K))
    assert np.all(np.abs(x.var(axis=0) - var) < 5 * var * np.sqrt(2.0 / K))


def test_ssor_sampler_mixes_slower_than_mgmc():
    """IACT of the Gibbs sampler far above the MGMC one at 32^3 (the
    reference's Gibbs / MGMC comparison, python/plot_summary.py:25-29)."""
    from gpu_common import make_problem
    op, _ = make_problem(3, (32, 32, 32), 4)
    N = op.lattice.Nvertex
    f = np.zeros(N)
    probes = np.random.default_rng(8).choice(N, 8, replace=False)
    taus = {}
    for name in ("ssor", "mgmc"):
        if name == "ssor":
            s = mg.SSORSampler(op, nchains=32, precision="fp32", seed=5)
        else:
            s = mg.MultigridMCSampler(op, params(4), nchains=32, precision="fp32", seed=5)
        s.fix_rhs(f)
        s.set_state(np.zeros((32, N)))
        s.set_probes(probes)
        s.run(100, record=False)
        obs = s.run(600)
        taus[name] = max(mg.tau_int_chains(obs[:, :, p])[0] for p in range(len(probes)))
    assert taus["mgmc"] < 2.0
    assert taus["ssor"] > 2.5 * taus["mgmc"]


def test_preconditioner_needs_small_coarsest_level():
    """A one-level SSOR sampler on a large lattice has no dense coarse factor;
    the multigrid preconditioner (always Cholesky on the coarsest level)
    reports it instead of failing on the device."""
    from gpu_common import make_problem
    op, _ = make_problem(3, (32, 32, 32), 0)
    s = mg.SSORSampler(op, nchains=1)
    x = np.zeros(op.lattice.Nvertex)
    s.apply(np.ones(op.lattice.Nvertex), x)
    assert np.all(np.isfinite(x)) and np.abs(x).max() > 0
    with pytest.raises(mg.MGMCError):
        s.coarse(np.zeros(op.lattice.Nvertex))
