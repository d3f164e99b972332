"""Differentiable posterior mean: a torch.autograd.Function over the C ABI.

For the shifted-Laplace posterior Q(theta) = h^d (-s Delta_h + kappa^2) +
B Sigma^{-1} B^T with Sigma = diag(sigma2), the posterior means of a batch of
observation vectors y_r (prior mean zero) are

    mu_r = Q^{-1} B Sigma^{-1} y_r.

Forward: one batched multigrid solve (FP32 or FP64 cycles, FP64 residuals and
iterative refinement). Backward: Q is symmetric, so the adjoint fields are a
second batched solve lam_r = Q^{-1} g_r with g_r = dL/dmu_r, and

    dL/dkappa^2 = -h^d sum_r lam_r . mu_r
    dL/ds       = -sum_r lam_r^T K mu_r,        K = h^d (-Delta_h)
    dL/dy_r     = Sigma^{-1} B^T lam_r
    dL/dsigma2_j = -sum_r (b_j^T lam_r)(y_rj - b_j^T mu_r) / sigma2_j^2

The two field contractions run in the HIP kernel k_contract (mgmc_contract);
the m-vectors B^T lam, B^T mu are sparse products on the host. Only the fine
fields mu are kept between forward and backward.
"""
import numpy as np
import torch

from .operators import MeasuredOperator, ShiftedLaplaceFDOperator
from .sampler import LoopSolver, MultigridParameters, MultigridPreconditioner


class PosteriorMeanProblem:
    """Fixed lattice and measurement geometry; the parameters kappa^2, s, y,
    sigma2 are the inputs of posterior_mean()."""

    def __init__(self, lattice, centres, radius, params=None, precision="fp32", rtol=1e-10, maxiter=100):
        self.lattice = lattice
        self.centres = np.asarray(centres, dtype=np.float64)
        self.radius = float(radius)
        self.params = params or MultigridParameters(nlevel=_nlevel(lattice))
        self.precision = precision
        self.rtol = rtol
        self.maxiter = maxiter
        m = len(self.centres)
        probe = MeasuredOperator(ShiftedLaplaceFDOperator(lattice, 1.0), self.centres, self.radius, np.ones(m))
        self.B = probe.get_B().tocsr()
        self.cell_volume = float(np.prod([1.0 / v for v in lattice.n]))

    def operator(self, kappa_sq, s, sigma2):
        prior = ShiftedLaplaceFDOperator(self.lattice, float(kappa_sq), diffusion=float(s))
        return MeasuredOperator(prior, self.centres, self.radius, np.asarray(sigma2, dtype=np.float64), B=self.B)

    def solve(self, op, rhs):
        """Q^{-1} rhs for each row; returns (solution, the handle used)."""
        rhs = np.atleast_2d(rhs)
        pre = MultigridPreconditioner(op, self.params, precision=self.precision, nchains=rhs.shape[0])
        solver = LoopSolver(op, pre, rtol=self.rtol, atol=1e300, maxiter=self.maxiter)
        x = np.asarray(solver.apply(rhs)).reshape(rhs.shape)
        return x, pre


def _nlevel(lattice):
    n = min(lattice.n)
    L = 1
    while n % 2 == 0 and n > 8:
        n //= 2
        L += 1
    return L


class _PosteriorMean(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa_sq, s, y, sigma2, problem):
        k, sv = float(kappa_sq), float(s)
        yy = y.detach().cpu().numpy().astype(np.float64)
        s2 = sigma2.detach().cpu().numpy().astype(np.float64)
        op = problem.operator(k, sv, s2)
        rhs = (problem.B @ (yy / s2).T).T
        mu, _ = problem.solve(op, rhs)
        ctx.problem = problem
        ctx.op = op
        ctx.mu = mu
        ctx.yy = yy
        ctx.s2 = s2
        return torch.from_numpy(mu.copy()).to(y.dtype)

    @staticmethod
    def backward(ctx, grad_mu):
        problem, op, mu, yy, s2 = ctx.problem, ctx.op, ctx.mu, ctx.yy, ctx.s2
        g = grad_mu.detach().cpu().numpy().astype(np.float64).reshape(mu.shape)
        lam, handle = problem.solve(op, g)
        dots = handle.contract(op.base.laplacian_stencil(), lam, mu)  # (lam.mu, lam^T K mu) per row
        d_kappa = -problem.cell_volume * dots[:, 0].sum()
        d_s = -dots[:, 1].sum()
        Btl = (problem.B.T @ lam.T).T  # (R, m)
        Btm = (problem.B.T @ mu.T).T
        d_y = Btl / s2
        d_sigma2 = -(Btl * (yy - Btm)).sum(axis=0) / (s2 * s2)
        t = grad_mu.dtype
        return (torch.tensor(d_kappa, dtype=t), torch.tensor(d_s, dtype=t), torch.from_numpy(d_y).to(t),
                torch.from_numpy(d_sigma2).to(t), None)


def posterior_mean(kappa_sq, s, y, sigma2, problem):
    """Posterior means (R, N) for observation batch y (R, m): differentiable
    with respect to kappa_sq, s (0-d tensors), y and sigma2 (m,)."""
    return _PosteriorMean.apply(kappa_sq, s, y, sigma2, problem)
