"""Full-problem Cholesky samplers (host): the reference's
SparseCholeskySampler, LowRankCholeskySampler and DenseCholeskySampler
(src/sampler/cholesky_sampler.cc:9-129, src/sampler/cholesky_sampler.hh:52-90).

Each draws independent samples x ~ N(Q^{-1} f, Q^{-1}) of the posterior
Q = A + B Sigma^{-1} B^T (or of the prior) through libmgmc_amd.so's host
factorisation: nested-dissection ordering, supernodal multifrontal Cholesky
on host threads, and blocked triangular solves for the chains. These are the
baselines the MGMC sampler is measured against; no GPU is used.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, f64, ptr
from .sampler import _problem

_FACTORISATION = {"sparse": 0, "lowrank": 1, "dense": 2}


class CholeskySampler:
    """CholeskySampler<LLT> (cholesky_sampler.hh:52-90) over nchains chains.

    apply(f, x) overwrites x (N, or nchains x N) with independent samples;
    fix_rhs(f) caches the solve with f as the reference does. nthreads = 0
    uses every hardware thread of the host."""

    factorisation = "sparse"

    def __init__(self, linear_operator, seed=0x5eed, nchains=1, nthreads=0, verbose=False):
        self.linear_operator = linear_operator
        self.N = linear_operator.lattice.Nvertex
        self.nchains = int(nchains)
        p, keep = _problem(linear_operator)
        h = C.c_void_p()
        check(_lib.lib().mgmc_cholesky_create(C.byref(p), _FACTORISATION[self.factorisation], self.nchains,
                                              int(seed), int(nthreads), C.byref(h)))
        self.h = h
        del keep
        n, nnz, ns, sec = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        check(_lib.lib().mgmc_cholesky_info(self.h, C.byref(n), C.byref(nnz), C.byref(ns), C.byref(sec)))
        self.nnz_L, self.nsupernodes, self.setup_seconds = nnz.value, ns.value, sec.value
        if verbose:
            # the print-out of the reference's factorisation wrappers
            # (src/auxilliary/cholesky_wrapper.cc:85-91)
            total = self.N * (self.N + 1) // 2
            print(f"==== {type(self).__name__} ====")
            print(f"  matrix size = {self.N} x {self.N}")
            print(f"  number of non-zero entries in L = nnz(L) = {self.nnz_L} of {total} "
                  f"[ {100.0 * self.nnz_L / total}% ]")
            print(f"  supernodes = {self.nsupernodes}, set-up time = {self.setup_seconds:.3f} s")

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().mgmc_cholesky_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_linear_operator(self):
        return self.linear_operator

    def sample(self, f=None, xi=None):
        """nchains x N independent samples; xi (nchains x N, optional) replaces
        the Philox normals (consumed in the factor's order)."""
        x = np.empty((self.nchains, self.N))
        ff = None if f is None else f64(f)
        per_chain = int(ff is not None and ff.size != self.N)
        xx = None if xi is None else np.ascontiguousarray(np.broadcast_to(f64(xi), (self.nchains, self.N)))
        check(_lib.lib().mgmc_cholesky_sample(self.h, ptr(ff), per_chain, ptr(xx), ptr(x)))
        return x

    def apply(self, f, x):
        """Sampler::apply(f, x): x receives a new sample per chain (its
        incoming value is not used)."""
        x[...] = self.sample(f).reshape(x.shape)

    def fix_rhs(self, f):
        check(_lib.lib().mgmc_cholesky_fix_rhs(self.h, ptr(f64(f))))

    def unfix_rhs(self):
        check(_lib.lib().mgmc_cholesky_unfix_rhs(self.h))


class SparseCholeskySampler(CholeskySampler):
    """Factorises Q = A + B Sigma^{-1} B^T (cholesky_sampler.cc:9-23)."""
    factorisation = "sparse"


class LowRankCholeskySampler(CholeskySampler):
    """Factorises the prior A and applies the QR-based low-rank correction
    for B Sigma^{-1} B^T (cholesky_sampler.cc:41-129)."""
    factorisation = "lowrank"


class DenseCholeskySampler(CholeskySampler):
    """Dense factorisation of Q in the reference's vertex order
    (cholesky_sampler.cc:26-38); at most 16384 unknowns."""
    factorisation = "dense"


def _time_samplers(argv=None):
    """Time per sample of the Cholesky samplers on n^3 lattices: the prior
    kappa^2 = 10 with seeded ball measurements (radius 0.05, variance 1e-4),
    right-hand side fixed, the reference's "time per sample" for the
    Cholesky column of its summary (python/plot_summary.py:31)."""
    import argparse
    import time

    from .operators import Lattice, MeasuredOperator, ShiftedLaplaceFDOperator, seeded_ball_measurements

    ap = argparse.ArgumentParser(prog="python -m multigridmc_amd.cholesky")
    ap.add_argument("sizes", type=int, nargs="+", help="cells per direction (3D lattices)")
    ap.add_argument("--factorisation", choices=sorted(_FACTORISATION), action="append")
    ap.add_argument("--measurements", type=int, default=32)
    ap.add_argument("--threads", type=int, default=0, help="host threads (0: all)")
    ap.add_argument("--samples", type=int, default=5)
    a = ap.parse_args(argv)
    classes = {c.factorisation: c for c in (SparseCholeskySampler, LowRankCholeskySampler, DenseCholeskySampler)}
    for n in a.sizes:
        lat = Lattice(n, n, n)
        centres, _, var = seeded_ball_measurements(lat, a.measurements, 0.05, 1e-4, 3)
        op = MeasuredOperator(ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.05, var)
        f = np.random.default_rng(1).standard_normal(lat.Nvertex)
        for name in a.factorisation or ["sparse", "lowrank"]:
            row = {}
            for K in (1, 8):
                s = classes[name](op, nchains=K, nthreads=a.threads)
                s.fix_rhs(f)
                s.sample()
                t0 = time.perf_counter()
                for _ in range(a.samples):
                    s.sample()
                row[K] = 1e3 * (time.perf_counter() - t0) / (a.samples * K)
                setup, nnz = s.setup_seconds, s.nnz_L
                s.close()
            print(f"{n}^3 {name:8s} N = {lat.Nvertex:8d}  nnz(L) = {nnz / 1e6:7.1f} M  set-up {setup:7.2f} s  "
                  f"ms/sample: 1 chain {row[1]:8.2f}, 8 chains {row[8]:8.2f}", flush=True)


if __name__ == "__main__":
    _time_samplers()
