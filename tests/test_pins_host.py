"""Known answers that pin the noise generator and the measurement vectors
(CPU only).

* Philox4x32-10: the known-answer vectors published with Random123
  (Salmon et al., SC'11; kat_vectors of the Random123 distribution) for the
  C++ header the kernels include (multigridmc_amd/csrc/philox.h, compiled
  here with g++) and for the NumPy restatement the GPU normals are checked
  against (tests/gpu_common.py).
* Ball-average measurement vectors (src/linear_operator/measured_operator.cc:69-170),
  for the oracle, the Python operator and the library's C entry point:
  - a vertex whose 2^d cells lie inside the ball gets exactly h^d / V_d(R)
    (the hat function integrates to h^d and the top-hat density is 1/V_d(R),
    V_2 = pi R^2, V_3 = 4/3 pi R^3);
  - zeroth and first moments: sum_j b_j and sum_j b_j x_j equal the 2^d-point
    Gauss-Legendre quadrature of the indicator and of x over the ball,
    enumerated here independently over all cells, because hat functions
    reproduce constants and linear functions. A wrong Gauss point, hat
    orientation, normalisation or skipped cell breaks one of them;
  - the textbook two-point rule these use (0.5 -+ 0.5/sqrt(3), weights 1/2)
    integrates x^k on [0, 1] exactly for k <= 3 (test_quadrature.hh:53-76).
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import multigridmc_amd as mg
from gpu_common import philox

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "multigridmc_amd", "csrc")

# (counter, key) -> output, philox4x32_10
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_header_known_answers(tmp_path):
    src = tmp_path / "kat.cc"
    lines = ['#include "philox.h"', "#include <cstdio>", "int main() {", "  int bad = 0;"]
    for ctr, key, want in KAT:
        lines.append("  { mgmc::u32x4 r = mgmc::philox4x32_10(%su, %su, %su, %su, %su, %su);" % (*ctr, *key))
        lines.append("    const unsigned w[4] = {%su, %su, %su, %su};" % want)
        lines.append("    for (int i = 0; i < 4; ++i) bad += r.v[i] != w[i]; }")
    lines += ['  std::printf("%d\\n", bad);', "  return bad != 0;", "}"]
    src.write_text("\n".join(lines))
    exe = str(tmp_path / "kat")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout


def test_numpy_philox_known_answers():
    for ctr, key, want in KAT:
        got = philox(*ctr, *key)
        assert [int(v) for v in got] == list(want)


def _ball_volume(R, dim):
    return math.pi * R * R if dim == 2 else 4.0 / 3.0 * math.pi * R ** 3


def _gauss_moments(n, x0, R):
    dim = len(n)
    h = 1.0 / np.array(n, dtype=float)
    g = np.array([0.5 - 0.5 / math.sqrt(3.0), 0.5 + 0.5 / math.sqrt(3.0)])
    axes = [(np.arange(n[d])[:, None] + g[None, :]).reshape(-1) * h[d] for d in range(dim)]
    P = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim)
    inside = np.linalg.norm(P - x0, axis=1) < R
    w = np.prod(h) / 2 ** dim / _ball_volume(R, dim)
    return w * inside.sum(), w * P[inside].sum(axis=0)


def _c_abi_vector(lat, x0, R):
    from multigridmc_amd import _lib
    dim = lat.dim
    nn = np.array(list(lat.n) + [1] * (3 - dim), dtype=np.int32)
    cap = 1 << 16
    idx, val, nnz = np.zeros(cap, np.int64), np.zeros(cap), C.c_int64()
    x0a = np.ascontiguousarray(x0, dtype=np.float64)
    _lib.check(_lib.lib().mgmc_measurement_vector(dim, nn.ctypes.data, x0a.ctypes.data, R, cap, idx.ctypes.data,
                                                  val.ctypes.data, C.byref(nnz)))
    return idx[:nnz.value], val[:nnz.value]


CASES = [((32, 32, 32), (0.43, 0.51, 0.47), 0.1), ((64, 64, 64), (0.3, 0.55, 0.61), 0.05),
         ((64, 32), (0.41, 0.52), 0.07), ((16, 16, 16), (0.5, 0.5, 0.5), 0.2), ((48, 40), (0.37, 0.61), 0.013)]


@pytest.mark.parametrize("n,x0,R", CASES)
@pytest.mark.parametrize("impl", ["oracle", "python", "c_abi"])
def test_measurement_vector_closed_forms(n, x0, R, impl, oracle):
    lat = mg.Lattice(*n)
    dim = lat.dim
    x0 = np.array(x0)
    if impl == "oracle":
        b = oracle.Operator.fd(dim, n, 1.0).measurement_vector(x0, R)
        idx = np.nonzero(b)[0]
        val = b[idx]
    elif impl == "python":
        idx, val = mg.measurement_vector(lat, x0, R)
    else:
        idx, val = _c_abi_vector(lat, x0, R)
    X = lat.vertex_coordinates(idx)
    h = 1.0 / np.array(n, dtype=float)
    # vertices whose surrounding cells lie inside the ball (all cell corners
    # within R of the centre)
    far = np.sqrt(np.sum((np.abs(X - x0) + h) ** 2, axis=1)) < R
    want = np.prod(h) / _ball_volume(R, dim)
    if far.any():
        assert np.abs(val[far] - want).max() <= 1e-13 * want
    m0, m1 = _gauss_moments(n, x0, R)
    assert abs(val.sum() - m0) <= 1e-13 * m0
    assert np.abs(val @ X - m1).max() <= 1e-13 * np.abs(m1).max()


def test_two_point_gauss_rule_is_exact_to_degree_three():
    g = np.array([0.5 - 0.5 / math.sqrt(3.0), 0.5 + 0.5 / math.sqrt(3.0)])
    for k in range(4):
        assert abs(0.5 * np.sum(g ** k) - 1.0 / (k + 1)) < 1e-15
    assert abs(0.5 * np.sum(g ** 4) - 0.2) > 1e-3


# ---- Gaussian-weighted balls (a beyond-reference option, pinned by its
# normalisation, its moments and the reduction to the top-hat) ----

@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("width", [0.2, 0.5, 0.99, 1.0, 3.0, 1e3])
def test_gaussian_ball_integral(dim, width):
    from scipy.integrate import quad
    from multigridmc_amd.operators import gaussian_ball_integral
    surface = 2 * math.pi if dim == 2 else 4 * math.pi
    ref = surface * quad(lambda u: u ** (dim - 1) * math.exp(-u * u / (2 * width * width)), 0.0, 1.0,
                         epsabs=0, epsrel=1e-13)[0]
    assert abs(gaussian_ball_integral(width, dim) - ref) <= 1e-12 * ref


def _gauss_moments_weighted(n, x0, R, width):
    from multigridmc_amd.operators import gaussian_ball_integral
    dim = len(n)
    h = 1.0 / np.array(n, dtype=float)
    g = np.array([0.5 - 0.5 / math.sqrt(3.0), 0.5 + 0.5 / math.sqrt(3.0)])
    axes = [(np.arange(n[d])[:, None] + g[None, :]).reshape(-1) * h[d] for d in range(dim)]
    P = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim)
    xi = np.linalg.norm(P - x0, axis=1) / R
    f = np.where(xi < 1.0, np.exp(-0.5 * xi * xi / width ** 2), 0.0)
    w = np.prod(h) / 2 ** dim / (R ** dim * gaussian_ball_integral(width, dim))
    return w * f.sum(), w * (f[:, None] * P).sum(axis=0)


@pytest.mark.parametrize("n,x0,R", CASES[:4])
def test_gaussian_measurement_vector_moments_and_reduction(n, x0, R):
    from multigridmc_amd import _lib
    lat = mg.Lattice(*n)
    x0 = np.array(x0)
    idx, val = mg.measurement_vector(lat, x0, R, profile="gaussian", width=0.4)
    X = lat.vertex_coordinates(idx)
    m0, m1 = _gauss_moments_weighted(n, x0, R, 0.4)
    assert abs(val.sum() - m0) <= 1e-13 * m0
    assert np.abs(val @ X - m1).max() <= 1e-13 * np.abs(m1).max()
    # the C entry point gives the same vector (exp may differ in the last ulp)
    nn = np.array(list(n) + [1] * (3 - len(n)), dtype=np.int32)
    cap = 1 << 16
    ci, cv, nnz = np.zeros(cap, np.int64), np.zeros(cap), C.c_int64()
    _lib.check(_lib.lib().mgmc_measurement_vector_weighted(lat.dim, nn.ctypes.data, x0.ctypes.data, R, 1, 0.4, cap,
                                                           ci.ctypes.data, cv.ctypes.data, C.byref(nnz)))
    assert np.array_equal(ci[:nnz.value], idx)
    assert np.abs(cv[:nnz.value] - val).max() <= 1e-15 * np.abs(val).max()
    # a very wide Gaussian is the top-hat
    it, vt = mg.measurement_vector(lat, x0, R)
    iw, vw = mg.measurement_vector(lat, x0, R, profile="gaussian", width=1e7)
    assert np.array_equal(it, iw) and np.abs(vw - vt).max() <= 1e-13 * np.abs(vt).max()


def test_gaussian_measurement_vector_normalised():
    """sum_j b_j -> 1 at second order as h -> 0 (ball away from the boundary)."""
    errs = []
    for n in (32, 64, 128):
        _, val = mg.measurement_vector(mg.Lattice(n, n, n), (0.43, 0.51, 0.47), 0.1, profile="gaussian", width=0.5)
        errs.append(abs(val.sum() - 1.0))
    assert errs[2] < 2e-3 and errs[2] < errs[0] / 4
