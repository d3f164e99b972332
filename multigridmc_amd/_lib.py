"""ctypes binding of libmgmc_amd.so (the C ABI declared in include/mgmc_amd.h).

The library is built in-tree by multigridmc_amd/build.py. There is no
fallback: if the shared object is missing or does not load, every entry point
raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmgmc_amd.so")

MGMC_OK = 0
STATUS = {1: "ERR_ARG", 2: "ERR_UNSUPPORTED", 3: "ERR_NOT_SPD", 4: "ERR_HIP", 5: "ERR_NOCONV", 6: "ERR_STATE"}


class MGMCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("n", C.c_int * 3),
        ("kappa_sq", C.c_double),
        ("A_rowptr", C.c_void_p),
        ("A_col", C.c_void_p),
        ("A_val", C.c_void_p),
        ("m", C.c_int),
        ("B_rowptr", C.c_void_p),
        ("B_col", C.c_void_p),
        ("B_val", C.c_void_p),
        ("sigma", C.c_void_p),
        ("diffusion", C.c_void_p),
        ("kappa_sq_field", C.c_void_p),
    ]


class Params(C.Structure):
    _fields_ = [
        ("nlevel", C.c_int),
        ("smoother", C.c_int),
        ("npresmooth", C.c_int),
        ("npostsmooth", C.c_int),
        ("ncoarsesmooth", C.c_int),
        ("omega", C.c_double),
        ("cycle", C.c_int),
        ("coarse_scaling", C.c_double),
        ("coarse_solver", C.c_int),
        ("precision", C.c_int),
        ("nchains", C.c_int),
        ("seed", C.c_uint64),
        ("stream", C.c_void_p),
    ]


# name -> (restype, argtypes); every function returns an int status
_vp = C.c_void_p
_SIGS = {
    "mgmc_create": [C.POINTER(Problem), C.POINTER(Params), C.POINTER(_vp)],
    "mgmc_destroy": [_vp],
    "mgmc_ndof": [_vp, C.c_int, C.POINTER(C.c_int64)],
    "mgmc_sample": [_vp, _vp, C.c_int, _vp],
    "mgmc_set_rhs": [_vp, _vp, C.c_int],
    "mgmc_set_state": [_vp, _vp],
    "mgmc_get_state": [_vp, _vp],
    "mgmc_set_probes": [_vp, C.c_int, _vp],
    "mgmc_set_observables": [_vp, C.c_int, _vp, _vp, _vp],
    "mgmc_run": [_vp, C.c_int, _vp, C.c_int],
    "mgmc_run_dev": [_vp, C.c_int, _vp, C.c_int],
    "mgmc_moments": [_vp, C.c_int, _vp, _vp],
    "mgmc_batch_create": [C.c_int, _vp, C.POINTER(_vp)],
    "mgmc_batch_destroy": [_vp],
    "mgmc_batch_run_dev": [_vp, C.c_int, _vp, C.c_int],
    "mgmc_batch_set_persistent": [_vp, C.c_int],
    "mgmc_batch_stats": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "mgmc_set_counter": [_vp, C.c_uint64],
    "mg_precondition": [_vp, _vp, _vp],
    "mg_solve": [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, _vp, C.POINTER(C.c_int)],
    "mgmc_apply_operator": [_vp, _vp, _vp],
    "mg_solve_dev": [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, _vp, C.POINTER(C.c_int)],
    "mgmc_measurement_dot_dev": [_vp, _vp, _vp],
    "mgmc_measurement_apply_dev": [_vp, _vp, _vp],
    "mgmc_sweep": [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp],
    "mgmc_residual_restrict": [_vp, C.c_int, _vp, _vp, _vp],
    "mgmc_prolongate_add": [_vp, C.c_int, C.c_double, _vp, _vp],
    "mgmc_coarse": [_vp, _vp, _vp, C.c_int, _vp],
    "mgmc_normals": [_vp, C.c_uint32, _vp],
    "mgmc_level_stencil": [_vp, C.c_int, _vp, C.POINTER(C.c_int)],
    "mgmc_level_stencils": [C.POINTER(Problem), C.c_int, C.c_int, _vp],
    "mgmc_measurement_vector": [C.c_int, _vp, _vp, C.c_double, C.c_int64, _vp, _vp, C.POINTER(C.c_int64)],
    "mgmc_fem_stencil": [C.c_int, _vp, C.c_double, _vp],
    "mgmc_measurement_vector_weighted": [C.c_int, _vp, _vp, C.c_double, C.c_int, C.c_double, C.c_int64, _vp, _vp,
                                         C.POINTER(C.c_int64)],
    "mgmc_contract": [_vp, _vp, _vp, _vp, _vp],
    "mgmc_contract_noalloc": [_vp, _vp, _vp, _vp, _vp],
    "mgmc_set_tuning": [_vp, C.c_char_p, C.c_int],
    "mgmc_set_ordering": [_vp, C.c_int],
    "mgmc_autotune": [_vp, C.c_int, C.POINTER(C.c_double)],
    "mgmc_get_tuning": [_vp, C.c_char_p, C.POINTER(C.c_int)],
    "mgmc_profile": [_vp, C.c_int],
    "mgmc_profile_result": [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "mgmc_cholesky_create": [C.POINTER(Problem), C.c_int, C.c_int, C.c_uint64, C.c_int, C.POINTER(_vp)],
    "mgmc_cholesky_destroy": [_vp],
    "mgmc_cholesky_info": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                           C.POINTER(C.c_double)],
    "mgmc_cholesky_set_nchains": [_vp, C.c_int],
    "mgmc_cholesky_fix_rhs": [_vp, _vp],
    "mgmc_cholesky_unfix_rhs": [_vp],
    "mgmc_cholesky_sample": [_vp, _vp, C.c_int, _vp, _vp],
}

_LIB = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m multigridmc_amd.build` (no CPU fallback exists)")
    # PyTorch (device tensors handed to the _dev entry points) ships its own
    # HIP runtime; loading it first makes this library bind to the same one,
    # so that torch still finds the GPU after the library has initialised it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.restype = C.c_int
        f.argtypes = args
    L.mgmc_last_error.restype = C.c_char_p
    L.mgmc_last_error.argtypes = []
    _LIB = L
    return L


def check(code):
    if code != MGMC_OK:
        raise MGMCError(code, lib().mgmc_last_error().decode())


def ptr(a):
    """Pointer to a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)
