"""ctypes binding of the FP64 CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_op_fd": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
        "orc_op_fd_periodic": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
        "orc_op_custom": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_long, _lp, _lp, _dp]),
        "orc_op_set_lowrank": (C.c_int, [vp, C.c_int, C.c_long, _lp, _lp, _dp, _dp]),
        "orc_op_add_measurements": (C.c_int, [vp, C.c_int, _dp, C.c_double, _dp, C.c_double, C.c_int, C.c_double]),
        "orc_op_free": (None, [vp]),
        "orc_op_ndof": (C.c_long, [vp]),
        "orc_op_m": (C.c_int, [vp]),
        "orc_op_nnz": (C.c_long, [vp, C.c_int]),
        "orc_op_csr": (None, [vp, C.c_int, _lp, _lp, _dp]),
        "orc_op_sigma": (None, [vp, _dp]),
        "orc_op_apply": (None, [vp, _dp, _dp]),
        "orc_op_dense_precision": (None, [vp, _dp]),
        "orc_op_coarsen": (vp, [vp]),
        "orc_restrict": (C.c_int, [vp, _dp, _dp]),
        "orc_prolongate_add": (C.c_int, [vp, C.c_double, _dp, _dp]),
        "orc_measurement_vector": (C.c_int, [vp, _dp, C.c_double, _dp]),
        "orc_op_dense_solve": (C.c_int, [vp, _dp, _dp]),
        "orc_mg_create": (vp, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                               C.c_double, C.c_int, C.c_int]),
        "orc_mg_free": (None, [vp]),
        "orc_mg_nlevel": (C.c_int, [vp]),
        "orc_mg_level_ndof": (C.c_long, [vp, C.c_int]),
        "orc_mg_level_nnz": (C.c_long, [vp, C.c_int, C.c_int]),
        "orc_mg_level_csr": (None, [vp, C.c_int, C.c_int, _lp, _lp, _dp]),
        "orc_mg_level_apply": (None, [vp, C.c_int, _dp, _dp]),
        "orc_mg_level_bbar": (None, [vp, C.c_int, C.c_int, _dp]),
        "orc_mg_sor_smoother": (None, [vp, C.c_int, C.c_int, _dp, _dp]),
        "orc_mg_sor_sampler_noise": (None, [vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
        "orc_mg_coarse_sample_noise": (None, [vp, _dp, _dp, _dp]),
        "orc_mg_coarse_solve": (None, [vp, _dp, _dp]),
        "orc_mg_restrict": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "orc_mg_prolongate_add": (C.c_int, [vp, C.c_int, C.c_double, _dp, _dp]),
        "orc_mg_precondition": (None, [vp, _dp, _dp]),
        "orc_loop_solve": (C.c_int, [vp, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp]),
        "orc_mgmc_reset_rng": (None, [vp]),
        "orc_mgmc_apply": (None, [vp, _dp, _dp]),
        "orc_mgmc_apply_noise": (C.c_long, [vp, _dp, _dp, C.c_void_p]),
        "orc_mgmc_chains": (C.c_double, [vp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _lp, C.c_void_p]),
        "orc_rng_create": (vp, [C.c_int, C.c_int]),
        "orc_rng_free": (None, [vp]),
        "orc_rng_ints": (None, [vp, C.c_long, _ip]),
        "orc_rng_uniform": (None, [vp, C.c_long, _dp]),
        "orc_rng_normal": (None, [vp, C.c_long, _dp]),
        "orc_tau_int_reference": (C.c_double, [_dp, C.c_long, C.c_int]),
        "orc_num_threads": (C.c_int, []),
        "orc_driver_mg_rhs": (None, [C.c_long, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _check(ptr):
    if not ptr:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())
    return ptr


def _csr_to_scipy(rp, ci, v, shape):
    import scipy.sparse as sp
    return sp.csr_matrix((v, ci, rp), shape=shape)


class Operator:
    """FP64 restatement of LinearOperator (Q = A + B Sigma^-1 B^T)."""

    def __init__(self, handle, dim, n):
        self.h = handle
        self.dim = dim
        self.n = tuple(n)

    @classmethod
    def fd(cls, dim, n, kappa_sq):
        n3 = list(n) + [1] * (3 - len(n))
        return cls(_check(lib().orc_op_fd(dim, n3[0], n3[1], n3[2], kappa_sq)), dim, n)

    @classmethod
    def fd_periodic(cls, dim, n, lam_min, lam_max):
        n3 = list(n) + [1] * (3 - len(n))
        return cls(_check(lib().orc_op_fd_periodic(dim, n3[0], n3[1], n3[2], lam_min, lam_max)), dim, n)

    @classmethod
    def custom(cls, dim, n, rows, cols, vals):
        n3 = list(n) + [1] * (3 - len(n))
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        vals = np.ascontiguousarray(vals, np.float64)
        return cls(_check(lib().orc_op_custom(dim, n3[0], n3[1], n3[2], len(rows), rows, cols, vals)), dim, n)

    def set_lowrank(self, m, rows, cols, vals, sigma):
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        vals = np.ascontiguousarray(vals, np.float64)
        sigma = np.ascontiguousarray(sigma, np.float64)
        if lib().orc_op_set_lowrank(self.h, m, len(rows), rows, cols, vals, sigma):
            raise RuntimeError(lib().orc_last_error().decode())

    def add_measurements(self, centres, radius, variance, variance_scaling=1.0, measure_global=False,
                         variance_global=0.0):
        centres = np.ascontiguousarray(centres, np.float64).reshape(-1)
        variance = np.ascontiguousarray(variance, np.float64)
        if lib().orc_op_add_measurements(self.h, len(variance), centres, radius, variance, variance_scaling,
                                         int(measure_global), variance_global):
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_op_free(self.h)
            self.h = None

    @property
    def N(self):
        return lib().orc_op_ndof(self.h)

    @property
    def m(self):
        return lib().orc_op_m(self.h)

    def csr(self, which=0):
        N = self.N
        nnz = lib().orc_op_nnz(self.h, which)
        nrow = N
        rp = np.zeros(nrow + 1, np.int64)
        ci = np.zeros(nnz, np.int64)
        v = np.zeros(nnz, np.float64)
        lib().orc_op_csr(self.h, which, rp, ci, v)
        ncol = N if which == 0 else self.m
        return _csr_to_scipy(rp, ci, v, (nrow, ncol))

    def sigma(self):
        s = np.zeros(self.m)
        lib().orc_op_sigma(self.h, s)
        return s

    def apply(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.N)
        lib().orc_op_apply(self.h, x, y)
        return y

    def dense_precision(self):
        Q = np.zeros((self.N, self.N))
        lib().orc_op_dense_precision(self.h, Q.reshape(-1))
        return Q

    def dense_solve(self, b):
        x = np.zeros(self.N)
        if lib().orc_op_dense_solve(self.h, np.ascontiguousarray(b, np.float64), x):
            raise RuntimeError("not SPD")
        return x

    def coarsen(self):
        cn = tuple(v // 2 for v in self.n)
        return Operator(_check(lib().orc_op_coarsen(self.h)), self.dim, cn)

    def coarse_ndof(self):
        N = 1
        for v in self.n:
            N *= v // 2 - 1
        return N

    def restrict(self, x):
        xc = np.zeros(self.coarse_ndof())
        if lib().orc_restrict(self.h, np.ascontiguousarray(x, np.float64), xc):
            raise RuntimeError(lib().orc_last_error().decode())
        return xc

    def prolongate_add(self, alpha, xc, x):
        x = np.array(x, np.float64)
        if lib().orc_prolongate_add(self.h, alpha, np.ascontiguousarray(xc, np.float64), x):
            raise RuntimeError(lib().orc_last_error().decode())
        return x

    def measurement_vector(self, x0, radius):
        out = np.zeros(self.N)
        lib().orc_measurement_vector(self.h, np.ascontiguousarray(x0, np.float64), radius, out)
        return out


SOR, SSOR = 0, 1
LEX, MULTICOLOUR = 0, 1


class Multigrid:
    """Hierarchy restating MultigridMCSampler / MultigridPreconditioner."""

    def __init__(self, op, nlevel, smoother=SSOR, npresmooth=1, npostsmooth=1, ncoarsesmooth=1, omega=1.0,
                 cycle=1, coarse_scaling=1.0, coarse_solver=0, ordering=LEX):
        self.op = op
        self.h = _check(lib().orc_mg_create(op.h, nlevel, smoother, npresmooth, npostsmooth, ncoarsesmooth,
                                            omega, cycle, coarse_scaling, coarse_solver, ordering))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_mg_free(self.h)
            self.h = None

    @property
    def nlevel(self):
        return lib().orc_mg_nlevel(self.h)

    def ndof(self, level):
        return lib().orc_mg_level_ndof(self.h, level)

    def level_csr(self, level, which=0):
        N = self.ndof(level)
        nnz = lib().orc_mg_level_nnz(self.h, level, which)
        rp = np.zeros(N + 1, np.int64)
        ci = np.zeros(nnz, np.int64)
        v = np.zeros(nnz)
        lib().orc_mg_level_csr(self.h, level, which, rp, ci, v)
        ncol = N if which == 0 else self.op.m
        return _csr_to_scipy(rp, ci, v, (N, ncol))

    def level_apply(self, level, x):
        y = np.zeros(self.ndof(level))
        lib().orc_mg_level_apply(self.h, level, np.ascontiguousarray(x, np.float64), y)
        return y

    def bbar(self, level, forward=True):
        out = np.zeros(self.ndof(level) * self.op.m)
        lib().orc_mg_level_bbar(self.h, level, int(forward), out)
        return out.reshape(self.ndof(level), self.op.m)

    def sor_smoother(self, level, forward, b, x):
        x = np.array(x, np.float64)
        lib().orc_mg_sor_smoother(self.h, level, int(forward), np.ascontiguousarray(b, np.float64), x)
        return x

    def sor_sampler_noise(self, level, forward, f, x, xi, xi_lr):
        x = np.array(x, np.float64)
        xi_lr = np.ascontiguousarray(xi_lr if xi_lr is not None else np.zeros(1), np.float64)
        lib().orc_mg_sor_sampler_noise(self.h, level, int(forward), np.ascontiguousarray(f, np.float64), x,
                                       np.ascontiguousarray(xi, np.float64), xi_lr)
        return x

    def coarse_sample_noise(self, f, xi):
        x = np.zeros(self.ndof(self.nlevel - 1))
        lib().orc_mg_coarse_sample_noise(self.h, np.ascontiguousarray(f, np.float64), x,
                                         np.ascontiguousarray(xi, np.float64))
        return x

    def coarse_solve(self, b):
        x = np.zeros(self.ndof(self.nlevel - 1))
        lib().orc_mg_coarse_solve(self.h, np.ascontiguousarray(b, np.float64), x)
        return x

    def restrict(self, level, x):
        xc = np.zeros(self.ndof(level + 1))
        lib().orc_mg_restrict(self.h, level, np.ascontiguousarray(x, np.float64), xc)
        return xc

    def prolongate_add(self, level, alpha, xc, x):
        x = np.array(x, np.float64)
        lib().orc_mg_prolongate_add(self.h, level, alpha, np.ascontiguousarray(xc, np.float64), x)
        return x

    def precondition(self, b):
        x = np.zeros(self.ndof(0))
        lib().orc_mg_precondition(self.h, np.ascontiguousarray(b, np.float64), x)
        return x

    def loop_solve(self, b, rtol=1e-10, atol=1e-10, maxiter=100):
        x = np.zeros(self.ndof(0))
        hist = np.zeros(maxiter)
        it = lib().orc_loop_solve(self.h, np.ascontiguousarray(b, np.float64), x, rtol, atol, maxiter, hist)
        n = it + 1 if it >= 0 else maxiter
        return x, hist[:n], it

    def reset_rng(self):
        lib().orc_mgmc_reset_rng(self.h)

    def mgmc_apply(self, f, x):
        x = np.array(x, np.float64)
        lib().orc_mgmc_apply(self.h, np.ascontiguousarray(f, np.float64), x)
        return x

    def mgmc_noise_count(self):
        N = self.ndof(0)
        return lib().orc_mgmc_apply_noise(self.h, np.zeros(N), np.zeros(N), None)

    def mgmc_apply_noise(self, f, x, xi):
        x = np.array(x, np.float64)
        xi = np.ascontiguousarray(xi, np.float64)
        lib().orc_mgmc_apply_noise(self.h, np.ascontiguousarray(f, np.float64), x, xi.ctypes.data)
        return x

    def chains(self, f, x, nsamples, nthreads=1, probes=None):
        """Independent CLCG chains; returns (x, obs[nsamples, nchains, nprobe], seconds)."""
        x = np.array(x, np.float64, order="C")
        nchains = x.shape[0]
        probes = np.ascontiguousarray(probes if probes is not None else np.zeros(0), np.int64)
        obs = np.zeros((nsamples, nchains, len(probes)))
        t = lib().orc_mgmc_chains(self.h, np.ascontiguousarray(f, np.float64), x.reshape(-1), nchains, nsamples,
                                  nthreads, len(probes), probes, obs.ctypes.data if len(probes) else None)
        return x, obs, t


class CLCG:
    def __init__(self, nstreams=1, stream=0):
        self.h = lib().orc_rng_create(nstreams, stream)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_rng_free(self.h)

    def ints(self, n):
        out = np.zeros(n, np.int32)
        lib().orc_rng_ints(self.h, n, out)
        return out

    def uniform(self, n):
        out = np.zeros(n)
        lib().orc_rng_uniform(self.h, n, out)
        return out

    def normal(self, n):
        out = np.zeros(n)
        lib().orc_rng_normal(self.h, n, out)
        return out


def driver_mg_rhs(n):
    """driver_mg's right-hand side: std::normal_distribution draws from std::mt19937_64(1482817)."""
    out = np.zeros(n)
    lib().orc_driver_mg_rhs(n, out)
    return out


def tau_int_reference(series, window):
    s = np.ascontiguousarray(series, np.float64)
    return lib().orc_tau_int_reference(s, len(s), window)
