"""Host-side operator setup mirroring the reference's lattice and operator classes.

Only setup lives here (NumPy/SciPy, FP64): lattices, the shifted-Laplace FD
prior and the ball-average measurement matrix B. Everything that runs per
sample is in the HIP library.
"""
import itertools
import math

import numpy as np
import scipy.sparse as sp


class Lattice:
    """Structured lattice of the unit square/cube with n[d] cells per direction;
    unknowns on the interior vertices, x fastest (src/lattice/lattice3d.hh:60-290)."""

    def __init__(self, *n):
        if len(n) not in (2, 3):
            raise ValueError("2D or 3D lattices only")
        self.n = tuple(int(v) for v in n)
        self.dim = len(n)

    @property
    def Nvertex(self):
        return int(np.prod([v - 1 for v in self.n]))

    @property
    def Ncell(self):
        return int(np.prod(self.n))

    def cell_volume(self):
        return float(np.prod([1.0 / v for v in self.n]))

    def vertex_euclidean(self, ell):
        """1-based Euclidean vertex indices of linear vertex indices (array)."""
        ell = np.asarray(ell, dtype=np.int64)
        out = []
        for v in self.n:
            out.append(ell % (v - 1) + 1)
            ell = ell // (v - 1)
        return np.stack(out, axis=-1)

    def vertex_coordinates(self, ell):
        return self.vertex_euclidean(ell) * (1.0 / np.array(self.n, dtype=np.float64))

    def vertex_index(self, idx):
        idx = np.asarray(idx, dtype=np.int64)
        ell = np.zeros(idx.shape[:-1], dtype=np.int64)
        stride = 1
        for d, v in enumerate(self.n):
            ell += (idx[..., d] - 1) * stride
            stride *= v - 1
        return ell

    def get_coarse_lattice(self):
        for v in self.n:
            if v % 2 != 0:
                raise ValueError("cannot coarsen lattice [one of the extents is odd]")
            if not v // 2 > 1:
                raise ValueError("cannot coarsen lattice [resulting lattice would have no interior vertices]")
        return Lattice(*[v // 2 for v in self.n])


def Lattice2d(nx, ny):
    return Lattice(nx, ny)


def Lattice3d(nx, ny, nz):
    return Lattice(nx, ny, nz)


class ShiftedLaplaceFDOperator:
    """Prior precision h^d (-div(a grad_h) + kappa^2) with Dirichlet elimination
    (src/linear_operator/shiftedlaplace_fd_operator.cc:9-57): the reference's
    operator for a = 1 and kappa^2 = 1 / Lambda^2 constant
    (src/linear_operator/correlationlength_model.hh:45-66) or per vertex (its
    correlation-length models, :83-105). The HIP library assembles the
    stencils itself from (lattice, kappa_sq, fields)."""

    m_lowrank = 0

    def __init__(self, lattice, kappa_sq, diffusion=1.0):
        """kappa_sq: a constant or one value per vertex (reference order).
        diffusion: a constant factor s on the Laplacian, h^d (-s Delta_h +
        kappa^2) (1 in the reference; other constants are passed to the
        library as a CSR matrix), or one value of a per cell (cell (i, j, k)
        at (k n_y + j) n_x + i): -div(a grad) with the coupling across an
        edge the mean of a over the cells sharing it (passed to the library
        as a field; per-vertex stencils on every level)."""
        self.lattice = lattice
        self.kappa_field = None
        self.diffusion_field = None
        if np.ndim(kappa_sq) == 0:
            self.kappa_sq = float(kappa_sq)
        else:
            self.kappa_field = np.ascontiguousarray(kappa_sq, dtype=np.float64).reshape(lattice.Nvertex)
            self.kappa_sq = 0.0
        if np.ndim(diffusion) == 0:
            self.diffusion = float(diffusion)
        else:
            ncell = int(np.prod(lattice.n))
            self.diffusion_field = np.ascontiguousarray(diffusion, dtype=np.float64).reshape(ncell)
            self.diffusion = 1.0

    def is_variable(self):
        return self.kappa_field is not None or self.diffusion_field is not None

    def get_ndof(self):
        return self.lattice.Nvertex

    def stencil(self):
        """(diagonal, off-diagonals per direction) in the reference's arithmetic."""
        cv = 1.0
        hinv2 = []
        for v in self.lattice.n:
            h = 1.0 / v
            hinv2.append(1.0 / (h * h))
            cv *= h
        s = self.diffusion
        diag = cv * self.kappa_sq
        for d in range(self.lattice.dim):
            diag += 2.0 * cv * (s * hinv2[d] if s != 1.0 else hinv2[d])
        return diag, [-cv * (s * h2 if s != 1.0 else h2) for h2 in hinv2]

    def laplacian_stencil(self):
        """3^d stencil (library index order) of h^d (-Delta_h): the derivative
        of the operator with respect to the diffusion factor."""
        d = self.lattice.dim
        cv = float(np.prod([1.0 / v for v in self.lattice.n]))
        st = np.zeros(27 if d == 3 else 9)
        centre = 13 if d == 3 else 4
        strides = [1, 3, 9]
        for k, v in enumerate(self.lattice.n):
            hinv2 = float(v * v)
            st[centre] += 2.0 * cv * hinv2
            st[centre - strides[k]] -= cv * hinv2
            st[centre + strides[k]] -= cv * hinv2
        return st

    def _edge_means(self, d):
        """E[v] = mean of a over the cells sharing the edge from vertex v to
        v + e_d, for v with v_d in 0..n_d-1 and the other components interior
        (array indexed [k, j, i] in 3D, [j, i] in 2D), summed in the library's
        order (cells with the slower components first)."""
        L = self.lattice
        dim = L.dim
        a = self.diffusion_field.reshape(tuple(L.n[::-1]))  # [k, j, i]
        ax = dim - 1 - d  # array axis of direction d
        out = None
        cnt = 0
        # the other directions take cells v_q - 1 and v_q (q != d), in order z, y, x
        shifts = [(0,) if q == ax else (0, 1) for q in range(dim)]
        for combo in itertools.product(*shifts):
            sl = tuple(slice(None) if q == ax else slice(combo[q], a.shape[q] - 1 + combo[q]) for q in range(dim))
            term = a[sl]
            out = term.copy() if out is None else out + term
            cnt += 1
        return out / cnt

    def _variable_sparse(self):
        L = self.lattice
        dim = L.dim
        N = L.Nvertex
        cv = 1.0
        hinv2 = []
        for v in L.n:
            h = 1.0 / v
            hinv2.append(1.0 / (h * h))
            cv *= h
        kap = self.kappa_field if self.kappa_field is not None else np.full(N, self.kappa_sq)
        diag = cv * kap
        idx = L.vertex_euclidean(np.arange(N))  # 1-based (i, j, k)
        rows, cols, vals = [], [], []
        for d in range(dim):
            if self.diffusion_field is None:
                am = ap = np.ones(N)
            else:
                E = self._edge_means(d)
                # E index: direction d runs over 0..n_d-1, the others over 1..n_q-1 (stored 0-based)
                pos = [idx[:, q] - 1 for q in range(dim)]
                pos_m = list(pos)
                pos_p = list(pos)
                pos_m[d] = idx[:, d] - 1
                pos_p[d] = idx[:, d]
                am = E[tuple(pos_m[::-1])]
                ap = E[tuple(pos_p[::-1])]
            diag = diag + cv * hinv2[d] * (am + ap)
            for s, coef in ((-1, -cv * hinv2[d] * am), (1, -cv * hinv2[d] * ap)):
                j = idx.copy()
                j[:, d] += s
                ok = (j[:, d] >= 1) & (j[:, d] <= L.n[d] - 1)
                rows.append(np.arange(N)[ok])
                cols.append(L.vertex_index(j[ok]))
                vals.append(coef[ok])
        rows.append(np.arange(N))
        cols.append(np.arange(N))
        vals.append(diag)
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
        A.sort_indices()
        return A

    def get_sparse(self):
        if self.is_variable():
            return self._variable_sparse()
        L = self.lattice
        N = L.Nvertex
        idx = L.vertex_euclidean(np.arange(N))
        diag, off = self.stencil()
        rows, cols, vals = [np.arange(N)], [np.arange(N)], [np.full(N, diag)]
        for d in range(L.dim):
            for s in (-1, 1):
                j = idx.copy()
                j[:, d] += s
                ok = (j[:, d] >= 1) & (j[:, d] <= L.n[d] - 1)
                rows.append(np.arange(N)[ok])
                cols.append(L.vertex_index(j[ok]))
                vals.append(np.full(ok.sum(), off[d]))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


class ShiftedLaplaceFEMOperator:
    """Prior precision of multilinear (Q1) finite elements for -Delta + kappa^2
    with Dirichlet elimination (src/linear_operator/shiftedlaplace_fem_operator.cc:9-141),
    constant kappa^2 = 1 / Lambda^2. Assembled as the reference does, cell by
    cell with 2^d-point Gauss-Legendre quadrature (exact for Q1); with constant
    coefficients every interior row holds the same 3^d stencil, which is what
    the HIP library reads back from the CSR matrix (a box-stencil fine level)."""

    m_lowrank = 0

    def __init__(self, lattice, kappa_sq):
        self.lattice = lattice
        self.kappa_sq = float(kappa_sq)

    def get_ndof(self):
        return self.lattice.Nvertex

    def local_matrix(self):
        """Element matrix M[alpha, beta] (alpha, beta in {0,1}^d, first
        component slowest) including the cell volume."""
        d = self.lattice.dim
        h = np.array([1.0 / v for v in self.lattice.n])
        hinv2 = 1.0 / (h * h)
        cv = float(np.prod(h))
        bits = np.array(list(itertools.product((0, 1), repeat=d)))
        g1 = np.array([0.5 * (1.0 - 1.0 / math.sqrt(3.0)), 0.5 * (1.0 + 1.0 / math.sqrt(3.0))])
        qp = g1[bits]                      # (nq, d) quadrature points
        qw = np.full(len(qp), 0.5 ** d)    # weights on [0, 1]^d
        # phi_alpha(x) = prod_d (x_d if alpha_d else 1 - x_d)
        fac = np.where(bits[:, None, :] == 1, qp[None, :, :], 1.0 - qp[None, :, :])  # (nb, nq, d)
        phi = np.prod(fac, axis=2)
        dfac = np.where(bits[:, None, :] == 1, 1.0, -1.0) * np.ones_like(fac)
        grad = np.empty_like(fac)
        for j in range(d):
            others = np.prod(np.delete(fac, j, axis=2), axis=2) if d > 1 else np.ones(fac.shape[:2])
            grad[:, :, j] = dfac[:, :, j] * others
        M = np.zeros((len(bits), len(bits)))
        for a in range(len(bits)):
            for b in range(len(bits)):
                val = 0.0
                for q in range(len(qp)):
                    val += (self.kappa_sq * phi[a, q] * phi[b, q] + np.dot(grad[a, q], hinv2 * grad[b, q])) * qw[q]
                M[a, b] = val * cv
        return M, bits

    def stencil(self):
        """3^d stencil (offsets in {-1,0,1}^d, first component slowest) of an
        interior row: the sum of the element entries of all cells holding
        both vertices."""
        M, bits = self.local_matrix()
        d = self.lattice.dim
        offs = list(itertools.product((-1, 0, 1), repeat=d))
        st = {}
        for s in offs:
            v = 0.0
            for a, alpha in enumerate(bits):
                beta = alpha + np.array(s)
                if np.all((beta >= 0) & (beta <= 1)):
                    b = int(np.dot(beta, 1 << np.arange(d)[::-1]))
                    v += M[a, b]
            st[s] = v
        return st

    def get_sparse(self):
        L = self.lattice
        N = L.Nvertex
        idx = L.vertex_euclidean(np.arange(N))
        rows, cols, vals = [], [], []
        # offsets are given with the first lattice component slowest
        for s, v in self.stencil().items():
            j = idx + np.array(s)
            ok = np.all((j >= 1) & (j <= np.array(L.n) - 1), axis=1)
            rows.append(np.arange(N)[ok])
            cols.append(L.vertex_index(j[ok]))
            vals.append(np.full(ok.sum(), v))
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
        A.sort_indices()
        return A


def V_sphere(radius, dim):
    """Volume of the d-ball (src/linear_operator/measured_operator.cc:52-66)."""
    if dim == 0:
        return 1.0
    if dim == 1:
        return 2.0 * radius
    return 2.0 * math.pi / dim * radius * radius * V_sphere(radius, dim - 2)


def periodic_kappa_sq(lattice, Lambda_min, Lambda_max):
    """kappa^2 = 1 / Lambda(x)^2 at every vertex for the reference's
    PeriodicCorrelationLengthModel (src/linear_operator/correlationlength_model.hh:83-105):
    Lambda(x) = Lambda_1 + Lambda_2 prod_d cos(pi x_d), Lambda_1,2 = (max +- min) / 2."""
    L1 = 0.5 * (Lambda_max + Lambda_min)
    L2 = 0.5 * (Lambda_max - Lambda_min)
    X = lattice.vertex_coordinates(np.arange(lattice.Nvertex))
    lam = np.full(lattice.Nvertex, L2)
    for d in range(lattice.dim):
        lam = lam * np.cos(math.pi * X[:, d])
    lam = lam + L1
    return 1.0 / (lam * lam)


def block_diffusion(lattice, blocks, low, high, seed):
    """A piecewise-constant diffusion coefficient: `blocks` (per direction)
    equal blocks of cells with values drawn uniformly from [low, high] by
    numpy's PCG64 with `seed`; one value per cell (cell (i, j, k) at
    (k n_y + j) n_x + i)."""
    dim = lattice.dim
    blocks = [blocks] * dim if np.ndim(blocks) == 0 else list(blocks)
    for d in range(dim):
        if lattice.n[d] % blocks[d]:
            raise ValueError("the blocks must divide the lattice")
    vals = np.random.default_rng(seed).uniform(low, high, size=tuple(blocks[::-1]))
    for d in range(dim):
        ax = dim - 1 - d
        vals = np.repeat(vals, lattice.n[d] // blocks[d], axis=ax)
    return np.ascontiguousarray(vals.reshape(-1))


def gaussian_ball_integral(width, dim):
    """Integral of exp(-|u|^2 / (2 width^2)) over the unit d-ball (d = 2, 3):
    2 pi w^2 (1 - e^{-1/(2w^2)}) in 2D; in 3D 4 pi int_0^1 u^2
    e^{-u^2/(2w^2)} du, in closed form for w < 1 and as its power series for
    w >= 1 (where the closed form cancels)."""
    w = float(width)
    a = 0.5 / (w * w)
    if dim == 2:
        return -2.0 * math.pi * w * w * math.expm1(-a)
    if w < 1.0:
        return 4.0 * math.pi * w * w * w * (math.sqrt(0.5 * math.pi) * math.erf(1.0 / (math.sqrt(2.0) * w))
                                            - math.exp(-a) / w)
    total, term = 0.0, 1.0
    for k in range(40):
        total += term / (2 * k + 3)
        term *= -a / (k + 1)
    return 4.0 * math.pi * total


PROFILES = {"tophat": 0, "gaussian": 1}


def measurement_vector(lattice, x0, radius, profile="tophat", width=0.5):
    """Dual-space ball average w_j = int f(x) phi_j(x) dx with a weight f of
    integral one over the ball of `radius` around x0, 2^d-point Gauss-Legendre
    per cell (src/linear_operator/measured_operator.cc:69-170); radius < 1e-12
    selects the nearest vertex (:74-90). Returns (vertex indices, values).

    profile "tophat": the reference's f = 1 / V_d(R); "gaussian": f =
    exp(-xi^2 / (2 width^2)) / (R^d c_d), xi = |x - x0| / R < 1, c_d its
    integral over the unit ball (a Gaussian-weighted ball average, no
    counterpart in the reference)."""
    if profile not in PROFILES:
        raise ValueError(f"unknown measurement profile '{profile}'")
    L = lattice
    dim = L.dim
    x0 = np.asarray(x0, dtype=np.float64)
    if radius < 1e-12:
        X = L.vertex_coordinates(np.arange(L.Nvertex))
        dist = np.sqrt(((X - x0) ** 2).sum(axis=1))
        return np.array([int(np.argmin(dist))]), np.array([1.0])
    h = np.array([1.0 / v for v in L.n])
    cv = L.cell_volume()
    if profile == "tophat":
        norm = 1.0 / V_sphere(radius, dim)
    else:
        norm = 1.0 / (radius ** dim * gaussian_ball_integral(width, dim))
    # corners / quadrature points: Cartesian products with the first component slowest
    bits = np.array([[(c >> (dim - 1 - d)) & 1 for d in range(dim)] for c in range(1 << dim)])
    p1 = np.array([0.5 * (-1.0 / math.sqrt(3.0) + 1.0), 0.5 * (1.0 / math.sqrt(3.0) + 1.0)])
    qp = p1[bits]  # (nq, dim)
    qw = 0.5 ** dim
    # candidate cells: a bounding box of the ball
    lo = [max(0, int(math.floor((x0[d] - radius) / h[d])) - 1) for d in range(dim)]
    hi = [min(L.n[d] - 1, int(math.ceil((x0[d] + radius) / h[d])) + 1) for d in range(dim)]
    axes = [np.arange(lo[d], hi[d] + 1) for d in range(dim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    cells = np.stack([g.reshape(-1) for g in grids], axis=1)  # x fastest = ascending linear cell index
    corners = h * (cells[:, None, :] + bits[None, :, :])
    cdist = np.sqrt(((corners - x0) ** 2).sum(axis=2))
    overlap = (cdist < radius).any(axis=1)
    cmin, cmax = corners.min(axis=1), corners.max(axis=1)
    overlap |= np.all((cmin <= x0) & (x0 <= cmax), axis=1)
    cells = cells[overlap]
    xq = h * (qp[None, :, :] + cells[:, None, :])  # (ncell, nq, dim)
    xi = np.sqrt(((xq - x0) ** 2).sum(axis=2)) / radius
    inside = xi < 1.0
    fq = np.ones_like(xi) if profile == "tophat" else np.exp(-0.5 * xi * xi / (width * width))
    rows, vals = [], []
    for a in range(1 << dim):
        vert = cells + bits[a]
        ok = np.all((vert >= 1) & (vert <= np.array(L.n) - 1), axis=1)
        # the weight times the corner's basis function, factors in dimension order
        phihat = fq.copy()
        for d in range(dim):
            phihat = phihat * np.where(bits[a, d] == 0, 1.0 - qp[:, d], qp[:, d])[None, :]
        contrib = np.where(inside, phihat * qw * cv * norm, 0.0)
        # sum over quadrature points in the reference's order
        local = np.zeros(len(cells))
        for q in range(contrib.shape[1]):
            local = local + contrib[:, q]
        rows.append(np.where(ok, L.vertex_index(np.where(ok[:, None], vert, 1)), -1))
        vals.append(local)
    # accumulate cell by cell, corner by corner (reference order)
    rows = np.stack(rows, axis=1).reshape(-1)
    vals = np.stack(vals, axis=1).reshape(-1)
    keep = rows >= 0
    rows, vals = rows[keep], vals[keep]
    uniq, inv = np.unique(rows, return_inverse=True)
    acc = np.zeros(len(uniq))
    np.add.at(acc, inv, vals)
    return uniq, acc


class MeasuredOperator:
    """Posterior precision Q = A + B Sigma^{-1} B^T with B the ball-average
    measurement matrix and Sigma = variance_scaling * variance
    (src/linear_operator/measured_operator.cc:9-49)."""

    def __init__(self, base, centres, radius, variance, variance_scaling=1.0, B=None, profile="tophat", width=0.5):
        """B: an already computed measurement matrix for these centres and
        radius (reused when only the variances or the prior change).
        profile / width: the ball weight of measurement_vector ("tophat", the
        reference's, or "gaussian")."""
        self.base = base
        self.lattice = base.lattice
        self.kappa_sq = base.kappa_sq
        centres = np.asarray(centres, dtype=np.float64).reshape(-1, self.lattice.dim)
        self.centres = centres
        self.radius = float(radius)
        m = len(centres)
        self.m_lowrank = m
        self.Sigma = variance_scaling * np.asarray(variance, dtype=np.float64).reshape(m)
        if B is not None:
            self.B = B
            return
        rows, cols, vals = [], [], []
        for k in range(m):
            r, v = measurement_vector(self.lattice, centres[k], radius, profile, width)
            rows.append(r)
            cols.append(np.full(len(r), k))
            vals.append(v)
        N = self.lattice.Nvertex
        if m:
            self.B = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                   shape=(N, m))
        else:
            self.B = sp.csr_matrix((N, 0))
        self.B.sort_indices()

    def get_ndof(self):
        return self.lattice.Nvertex

    def get_sparse(self):
        return self.base.get_sparse()

    def get_B(self):
        return self.B

    def get_Sigma(self):
        return self.Sigma

    def apply(self, x):
        A = self.get_sparse()
        return A @ x + self.B @ ((self.B.T @ x) / self.Sigma)


def seeded_ball_measurements(lattice, m, radius, variance, seed, low=0.1, high=0.9):
    """m ball measurements with centres ~ U[low, high]^d and observed values
    y ~ N(0, 1) from numpy's PCG64 with the given seed."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(low, high, size=(m, lattice.dim))
    y = rng.standard_normal(m)
    var = np.full(m, float(variance))
    return centres, y, var
