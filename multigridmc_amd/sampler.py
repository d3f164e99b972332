"""Python host mirror of the reference's Sampler / Preconditioner / LinearSolver
classes, each a thin wrapper over the C ABI of libmgmc_amd.so.

Names, argument meaning and error behaviour follow the reference
(src/sampler/sampler.hh:23-78, src/sampler/multigridmc_sampler.hh,
src/preconditioner/multigrid_preconditioner.hh, src/solver/loop_solver.hh);
errors raise MGMCError instead of calling exit().
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, f64, i64, ptr
from .operators import ShiftedLaplaceFDOperator


@dataclass
class MultigridParameters:
    """MultigridParameters (src/auxilliary/parameters.hh:150-179)."""

    nlevel: int = 4
    smoother: str = "SSOR"
    coarse_solver: str = "Cholesky"
    npresmooth: int = 1
    npostsmooth: int = 1
    ncoarsesmooth: int = 1
    omega: float = 1.0
    cycle: int = 1
    coarse_scaling: float = 1.0
    verbose: int = 0


def _problem(op):
    """mgmc_problem for a prior operator (ShiftedLaplaceFDOperator: assembled
    by the library from kappa^2; any other constant-coefficient operator, e.g.
    ShiftedLaplaceFEMOperator: passed as its CSR matrix) or a MeasuredOperator
    wrapping one."""
    lat = op.lattice
    p = _lib.Problem()
    p.dim = lat.dim
    for d in range(3):
        p.n[d] = lat.n[d] if d < lat.dim else 1
    p.kappa_sq = op.kappa_sq
    keep = []
    prior = getattr(op, "base", op)
    if isinstance(prior, ShiftedLaplaceFDOperator) and prior.is_variable():
        # coefficient fields: the library assembles per-vertex stencils
        if prior.diffusion_field is not None:
            keep.append(prior.diffusion_field)
            p.diffusion = ptr(prior.diffusion_field)
        if prior.kappa_field is not None:
            keep.append(prior.kappa_field)
            p.kappa_sq_field = ptr(prior.kappa_field)
    elif not isinstance(prior, ShiftedLaplaceFDOperator) or prior.diffusion != 1.0:
        A = prior.get_sparse().tocsr()
        A.sort_indices()
        rp, ci, v = i64(A.indptr), i64(A.indices), f64(A.data)
        keep += [rp, ci, v]
        p.A_rowptr, p.A_col, p.A_val = ptr(rp), ptr(ci), ptr(v)
    m = getattr(op, "m_lowrank", 0)
    p.m = m
    if m:
        B = op.get_B().tocsr()
        B.sort_indices()
        rp, ci, v, s = i64(B.indptr), i64(B.indices), f64(B.data), f64(op.get_Sigma())
        keep += [rp, ci, v, s]
        p.B_rowptr, p.B_col, p.B_val, p.sigma = ptr(rp), ptr(ci), ptr(v), ptr(s)
    return p, keep


_SMOOTHERS = {"SOR": 0, "SSOR": 1}
_COARSE = {"Cholesky": 0, "SSOR": 1}
_PREC = {"fp64": 0, "fp32": 1}


_ORDERING = {"multicolour": 0, "lexicographic": 1}


class _Handle:
    def __init__(self, op, params, nchains, precision, seed, stream=None, ordering="multicolour"):
        if params.smoother not in _SMOOTHERS:
            raise _lib.MGMCError(1, f"invalid smoother '{params.smoother}'")
        if params.coarse_solver not in _COARSE:
            raise _lib.MGMCError(1, f"invalid coarse solver '{params.coarse_solver}'")
        if ordering not in _ORDERING:
            raise _lib.MGMCError(1, f"invalid ordering '{ordering}'")
        L = _lib.lib()
        prob, self._keep = _problem(op)
        par = _lib.Params()
        par.nlevel = params.nlevel
        par.smoother = _SMOOTHERS[params.smoother]
        par.npresmooth = params.npresmooth
        par.npostsmooth = params.npostsmooth
        par.ncoarsesmooth = params.ncoarsesmooth
        par.omega = params.omega
        par.cycle = params.cycle
        par.coarse_scaling = params.coarse_scaling
        par.coarse_solver = _COARSE[params.coarse_solver]
        par.precision = _PREC[precision]
        par.nchains = nchains
        par.seed = seed
        par.stream = stream
        h = C.c_void_p()
        check(L.mgmc_create(C.byref(prob), C.byref(par), C.byref(h)))
        self.h = h
        if ordering != "multicolour":
            check(L.mgmc_set_ordering(h, _ORDERING[ordering]))
        self.ordering = ordering
        self.nchains = nchains
        self.precision = precision
        self.op = op
        self.params = params
        self.N = op.lattice.Nvertex
        self.m = getattr(op, "m_lowrank", 0)

    def close(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.mgmc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_tuning(self, key, value):
        check(_lib.lib().mgmc_set_tuning(self.h, key.encode(), int(value)))

    def get_tuning(self, key):
        v = C.c_int()
        check(_lib.lib().mgmc_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    def autotune(self, ncycles=5):
        """Pick the fastest launch shapes for this handle (mgmc_autotune);
        returns the best time per cycle in ms. The chain state is kept."""
        ms = C.c_double()
        check(_lib.lib().mgmc_autotune(self.h, int(ncycles), C.byref(ms)))
        return ms.value

    def apply_operator(self, x):
        """y = A x + B Sigma^{-1} B^T x on the GPU for every row of x
        (LinearOperator::apply, src/linear_operator/linear_operator.hh:78-88)."""
        xx = self._rows(x)
        out = np.zeros_like(xx)
        check(_lib.lib().mgmc_apply_operator(self.h, ptr(xx), ptr(out)))
        return out.reshape(np.shape(x)) if np.size(x) == out.size else out

    def contract(self, stencil, lam, mu):
        """Per chain (lam . mu, lam^T S mu) for a constant stencil S
        (mgmc_contract): the adjoint contraction of parameter gradients."""
        ll, mm = self._rows(lam), self._rows(mu)
        out = np.zeros((self.nchains, 2))
        check(_lib.lib().mgmc_contract(self.h, ptr(f64(stencil)), ptr(ll), ptr(mm), ptr(out)))
        return out

    def solve_dev(self, b_dev, x_dev, rtol, atol, maxiter):
        """mg_solve_dev on device pointers (nchains x N each); returns the
        converged iteration (-1 if maxiter was reached without converging)."""
        it = C.c_int(-1)
        code = _lib.lib().mg_solve_dev(self.h, b_dev, x_dev, rtol, atol, maxiter, None, C.byref(it))
        if code != 5:
            check(code)
        return it.value

    def measurement_dot_dev(self, x_dev, z_dev):
        """z = B^T x for device pointers (nchains x N -> nchains x m)."""
        check(_lib.lib().mgmc_measurement_dot_dev(self.h, x_dev, z_dev))

    def measurement_apply_dev(self, w_dev, y_dev):
        """y = B w for device pointers (nchains x m -> nchains x N)."""
        check(_lib.lib().mgmc_measurement_apply_dev(self.h, w_dev, y_dev))

    def contract_noalloc(self, stencil, lam_dev, mu_dev, out_dev):
        """mgmc_contract_noalloc on device pointers (nchains x N reference-order
        FP64 inputs, 2 x nchains output): allocates nothing and does not
        synchronise the handle's stream."""
        check(_lib.lib().mgmc_contract_noalloc(self.h, ptr(f64(stencil)), lam_dev, mu_dev, out_dev))

    def ndof(self, level):
        n = C.c_int64()
        check(_lib.lib().mgmc_ndof(self.h, level, C.byref(n)))
        return n.value

    def _rows(self, a, level=0):
        n = self.ndof(level)
        a = f64(a)
        if a.size == n:
            a = np.ascontiguousarray(np.broadcast_to(a.reshape(1, n), (self.nchains, n)))
        return a.reshape(self.nchains, n)


class MultigridMCSampler(_Handle):
    """MGMC sampler (src/sampler/multigridmc_sampler.cc:8-145) for K batched chains.

    apply(f, x) draws one new sample per chain: x has shape (nchains, N) (or (N,)
    for a single chain) and is updated in place, as Sampler::apply does.
    ordering="lexicographic" sweeps in the reference's own vertex order
    (wavefronts, mgmc_set_ordering) instead of the fast multicolour order."""

    def __init__(self, linear_operator, params=None, nchains=1, precision="fp64", seed=0x5eed, stream=None,
                 ordering="multicolour"):
        super().__init__(linear_operator, params or MultigridParameters(), nchains, precision, seed, stream, ordering)

    def get_linear_operator(self):
        return self.op

    def apply(self, f, x):
        shape = x.shape
        xx = self._rows(x)
        ff = f64(f)
        per_chain = int(ff.size != self.N)
        check(_lib.lib().mgmc_sample(self.h, ptr(ff), per_chain, ptr(xx)))
        x[...] = xx.reshape(shape)

    def fix_rhs(self, f):
        ff = f64(f)
        check(_lib.lib().mgmc_set_rhs(self.h, ptr(ff), int(ff.size != self.N)))

    def unfix_rhs(self):
        pass

    def set_state(self, x):
        check(_lib.lib().mgmc_set_state(self.h, ptr(self._rows(x))))

    def get_state(self):
        out = np.zeros((self.nchains, self.N))
        check(_lib.lib().mgmc_get_state(self.h, ptr(out)))
        return out

    def set_probes(self, vertices):
        v = i64(vertices)
        check(_lib.lib().mgmc_set_probes(self.h, len(v), ptr(v)))
        self.nprobe = len(v)

    def run(self, nsamples, record=True, use_graph=True):
        """nsamples resident cycles; returns obs[nsamples, nchains, nprobe] or None."""
        obs = np.zeros((nsamples, self.nchains, getattr(self, "nprobe", 0))) if record else None
        check(_lib.lib().mgmc_run(self.h, nsamples, ptr(obs) if record and obs.size else None, int(use_graph)))
        return obs

    def moments(self, nsamples):
        """Pointwise mean and variance over `nsamples` further cycles of all
        chains (posterior_statistics, src/measurements/posterior_statistics.cc)."""
        mean, var = np.zeros(self.N), np.zeros(self.N)
        check(_lib.lib().mgmc_moments(self.h, nsamples, ptr(mean), ptr(var)))
        return mean, var

    def run_dev(self, nsamples, obs_dev_ptr=None, use_graph=True):
        check(_lib.lib().mgmc_run_dev(self.h, nsamples, obs_dev_ptr, int(use_graph)))

    def profile(self, enable):
        check(_lib.lib().mgmc_profile(self.h, int(enable)))

    def profile_result(self):
        ms, n = C.c_double(), C.c_int64()
        check(_lib.lib().mgmc_profile_result(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_counter(self, c):
        check(_lib.lib().mgmc_set_counter(self.h, c))

    # ---- component entry points ----
    def sweep(self, level, forward, stochastic, b, x, xi=None, xi_lr=None):
        bb, xx = self._rows(b, level), self._rows(x, level).copy()
        xi = None if xi is None else self._rows(xi, level)
        xl = None if xi_lr is None else f64(xi_lr).reshape(self.nchains, -1)
        check(_lib.lib().mgmc_sweep(self.h, level, int(forward), int(stochastic), ptr(bb), ptr(xx), ptr(xi),
                                    ptr(xl)))
        return xx

    def residual_restrict(self, level, f, x):
        out = np.zeros((self.nchains, self.ndof(level + 1)))
        check(_lib.lib().mgmc_residual_restrict(self.h, level, ptr(self._rows(f, level)), ptr(self._rows(x, level)),
                                                ptr(out)))
        return out

    def prolongate_add(self, level, alpha, xc, x):
        xx = self._rows(x, level).copy()
        check(_lib.lib().mgmc_prolongate_add(self.h, level, alpha, ptr(self._rows(xc, level + 1)), ptr(xx)))
        return xx

    def coarse(self, f, xi_mode=0, xi=None):
        lc = self.params.nlevel - 1
        out = np.zeros((self.nchains, self.ndof(lc)))
        xi = None if xi is None else self._rows(xi, lc)
        check(_lib.lib().mgmc_coarse(self.h, ptr(self._rows(f, lc)), ptr(out), xi_mode, ptr(xi)))
        return out

    def normals(self, tag):
        out = np.zeros((self.nchains, self.N))
        check(_lib.lib().mgmc_normals(self.h, tag, ptr(out)))
        return out

    def level_stencil(self, level):
        dim = self.op.lattice.dim
        st = np.zeros(27 if dim == 3 else 9)
        nc = C.c_int()
        check(_lib.lib().mgmc_level_stencil(self.h, level, ptr(st), C.byref(nc)))
        return st, nc.value

    def set_observables(self, observables):
        """Observables z_p = b_p . x, each given as (vertex indices, weights)
        (e.g. measurement_vector(lattice, x0, radius)); run() then records
        obs[sample, chain, p]."""
        rp = [0]
        idx, w = [], []
        for v, wt in observables:
            idx.append(i64(v))
            w.append(f64(wt))
            rp.append(rp[-1] + len(v))
        rp = i64(rp)
        vi = i64(np.concatenate(idx)) if idx else i64([])
        wv = f64(np.concatenate(w)) if w else f64([])
        check(_lib.lib().mgmc_set_observables(self.h, len(observables), ptr(rp), ptr(vi), ptr(wv)))
        self.nprobe = len(observables)


class SSORSampler(MultigridMCSampler):
    """Standalone SSOR Gibbs sampler (src/sampler/ssor_sampler.cc:9-15): each
    apply is `nsmooth` x (forward SOR sweep, backward SOR sweep) with noise
    and the low-rank correction, on the fine lattice only. Runs the same
    kernels as the multigrid cycle (a one-level hierarchy whose level is
    sampled by SSOR sweeps)."""

    def __init__(self, linear_operator, omega=1.0, nsmooth=1, nchains=1, precision="fp64", seed=0x5eed, stream=None):
        params = MultigridParameters(nlevel=1, smoother="SSOR", npresmooth=0, npostsmooth=0, ncoarsesmooth=nsmooth,
                                     omega=omega, coarse_solver="SSOR")
        super().__init__(linear_operator, params, nchains, precision, seed, stream)


class MultigridPreconditioner(_Handle):
    """Zero-initial-guess multigrid cycle (src/preconditioner/multigrid_preconditioner.cc:8-109)."""

    def __init__(self, linear_operator, params=None, precision="fp64", nchains=1, stream=None, ordering="multicolour"):
        super().__init__(linear_operator, params or MultigridParameters(), nchains, precision, 0, stream, ordering)

    def apply(self, b, x=None):
        bb = self._rows(b)
        out = np.zeros_like(bb)
        check(_lib.lib().mg_precondition(self.h, ptr(bb), ptr(out)))
        out = out.reshape(np.shape(b)) if np.size(b) == out.size else out
        if x is not None:
            x[...] = out
        return out


class LoopSolver:
    """Richardson iteration x -= P(Ax - b) (src/solver/loop_solver.cc:9-54),
    residuals in FP64 on the GPU whatever the cycle's precision."""

    def __init__(self, linear_operator, preconditioner, rtol=1e-10, atol=1e-10, maxiter=100, verbose=0):
        self.op = linear_operator
        self.pre = preconditioner
        self.rtol, self.atol, self.maxiter, self.verbose = rtol, atol, maxiter, verbose
        self.history = None
        self.niter = -1

    def apply(self, b, x=None):
        bb = self.pre._rows(b)
        out = np.zeros_like(bb)
        hist = np.zeros(self.maxiter)
        it = C.c_int(-1)
        code = _lib.lib().mg_solve(self.pre.h, ptr(bb), ptr(out), self.rtol, self.atol, self.maxiter, ptr(hist),
                                   C.byref(it))
        self.niter = it.value
        n = it.value + 1 if it.value >= 0 else self.maxiter
        self.history = hist[:n]
        if self.verbose >= 2:
            print("%5s   %8s   %12s   %6s" % ("iter", "||r||", "||r||/||r_0||", "rho"))
            r0 = np.linalg.norm(f64(b).reshape(-1)[: self.pre.N])
            old = r0
            for k, r in enumerate(self.history):
                print("%5d   %8.3e   %12.3e   %6.3f" % (k, r, r / r0, r / old))
                old = r
        if code == 5:
            if self.verbose >= 1:
                print("Solver failed to converge after %5d iterations" % self.maxiter)
        else:
            check(code)
        out = out.reshape(np.shape(b)) if np.size(b) == out.size else out
        if x is not None:
            x[...] = out
        return out
