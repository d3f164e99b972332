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

Everything stays in device memory: B Sigma^{-1} y, the solves, B^T lam and
B^T mu run on the library's device entry points (mgmc_measurement_apply_dev,
mg_solve_dev, mgmc_measurement_dot_dev) and the two field contractions in the
HIP kernel of mgmc_contract_noalloc; torch holds the tensors. Only the fine
fields mu and the handle built in the forward pass are kept for the backward
pass (the hierarchy is not rebuilt). Tensors on the CPU are moved to the GPU
and the results returned on the caller's device.
"""
import numpy as np
import torch

from .operators import MeasuredOperator, ShiftedLaplaceFDOperator
from .sampler import MultigridParameters, MultigridPreconditioner


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

    def solve_dev(self, handle, rhs, x):
        """x = Q^{-1} rhs row by row for device tensors (R x N, FP64) with the
        handle's multigrid loop solver (stopping on the relative residual)."""
        it = handle.solve_dev(rhs.data_ptr(), x.data_ptr(), self.rtol, 1e300, self.maxiter)
        if it < 0:
            raise RuntimeError(f"posterior-mean solve did not reach rtol {self.rtol} in {self.maxiter} iterations")
        return x


def _nlevel(lattice):
    n = min(lattice.n)
    L = 1
    while n % 2 == 0 and n > 8:
        n //= 2
        L += 1
    return L


def _ptr(t):
    return t.data_ptr()


class _PosteriorMean(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa_sq, s, y, sigma2, problem):
        dev = torch.device("cuda")
        k, sv = float(kappa_sq), float(s)
        # Sigma enters the hierarchy set-up on the host (m values)
        s2_host = sigma2.detach().to("cpu", torch.float64).numpy()
        op = problem.operator(k, sv, s2_host)
        yd = y.detach().to(dev, torch.float64).contiguous()
        s2d = sigma2.detach().to(dev, torch.float64).contiguous()
        R = yd.shape[0]
        N = problem.lattice.Nvertex
        h = MultigridPreconditioner(op, problem.params, precision=problem.precision, nchains=R)
        w = (yd / s2d).contiguous()
        rhs = torch.empty((R, N), dtype=torch.float64, device=dev)
        mu = torch.empty((R, N), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        h.measurement_apply_dev(_ptr(w), _ptr(rhs))
        problem.solve_dev(h, rhs, mu)
        ctx.problem, ctx.handle, ctx.op = problem, h, op
        ctx.save_for_backward(mu, yd, s2d)
        ctx.in_device = y.device
        return mu.to(y.device, y.dtype)

    @staticmethod
    def backward(ctx, grad_mu):
        problem, h, op = ctx.problem, ctx.handle, ctx.op
        mu, yd, s2d = ctx.saved_tensors
        dev = mu.device
        R, m = yd.shape
        g = grad_mu.detach().to(dev, torch.float64).contiguous()
        lam = torch.empty_like(mu)
        dots = torch.empty(2 * R, dtype=torch.float64, device=dev)
        Btl = torch.empty((R, m), dtype=torch.float64, device=dev)
        Btm = torch.empty((R, m), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        problem.solve_dev(h, g, lam)
        h.contract_noalloc(op.base.laplacian_stencil(), _ptr(lam), _ptr(mu), _ptr(dots))  # (lam.mu, lam^T K mu)
        h.measurement_dot_dev(_ptr(lam), _ptr(Btl))  # synchronises the handle's stream
        h.measurement_dot_dev(_ptr(mu), _ptr(Btm))
        d_kappa = -problem.cell_volume * dots[0::2].sum()
        d_s = -dots[1::2].sum()
        d_y = Btl / s2d
        d_sigma2 = -(Btl * (yd - Btm)).sum(dim=0) / (s2d * s2d)
        t, out = grad_mu.dtype, ctx.in_device
        return (d_kappa.to(out, t), d_s.to(out, t), d_y.to(out, t), d_sigma2.to(out, t), None)


def posterior_mean(kappa_sq, s, y, sigma2, problem):
    """Posterior means (R, N) for observation batch y (R, m): differentiable
    with respect to kappa_sq, s (0-d tensors), y and sigma2 (m,)."""
    return _PosteriorMean.apply(kappa_sq, s, y, sigma2, problem)
