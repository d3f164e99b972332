// Host-side (FP64) construction of the multigrid hierarchy.
//
// Reference behaviour reproduced here (paths relative to the reference root):
//  * fine operator: ShiftedLaplaceFDOperator (src/linear_operator/
//    shiftedlaplace_fd_operator.cc:9-57) or a caller-supplied CSR matrix,
//    recognised as a constant nearest-neighbour stencil,
//  * Galerkin coarsening A_c = R A R^T, B_c = R B, Sigma_c = Sigma
//    (src/linear_operator/linear_operator.cc:10-23) with R = P^T and P the
//    tensor product of {1/2, 1, 1/2} (src/intergrid/intergrid_operator_linear.cc:8-31),
//  * the SMW low-rank correction of SORSmoother (src/smoother/sor_smoother.cc:17-37),
//    kept factored as W = (L + D/omega)^{-1} B (sparse) and S^{-1} = (Sigma + B^T W)^{-1}
//    instead of the dense N x m matrix B W S^{-1},
//  * the coarsest-level Cholesky factor (src/sampler/cholesky_sampler.cc:9-23,
//    src/solver/cholesky_solver.cc:8-27), stored as explicit L^{-T} and Q^{-1}
//    so that the per-cycle triangular solves become dense products.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>

#include "../../include/mgmc_amd.h"

namespace mgmc {

static int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

Geom Geom::make(int dim, const int* n) {
  if (dim != 2 && dim != 3) throw Error(MGMC_ERR_UNSUPPORTED, "only dim = 2 or 3 is supported");
  Geom g;
  g.dim = dim;
  for (int d = 0; d < 3; ++d) g.n[d] = (d < dim) ? n[d] : 1;
  for (int d = 0; d < dim; ++d)
    if (g.n[d] < 2) throw Error(MGMC_ERR_ARG, "lattice extent must be >= 2 cells");
  g.px = (int)round_up(g.n[0], 8);
  g.py = g.n[1] + 1;
  g.pz = dim == 3 ? g.n[2] + 1 : 1;
  g.vol = (int64_t)g.px * g.py * g.pz;
  g.cstride = round_up(g.vol, 64);
  g.N = (int64_t)(g.n[0] - 1) * (g.n[1] - 1) * (dim == 3 ? g.n[2] - 1 : 1);
  if (g.vol >= (int64_t(1) << 31)) throw Error(MGMC_ERR_UNSUPPORTED, "level too large for 32-bit row ids");
  return g;
}

// Lattice3d::get_coarse_lattice (src/lattice/lattice3d.hh:248-263)
Geom Geom::coarse() const {
  int nc[3];
  for (int d = 0; d < dim; ++d) {
    if (n[d] % 2 != 0) throw Error(MGMC_ERR_ARG, "cannot coarsen lattice [one of the extents is odd]");
    if (!(n[d] / 2 > 1)) throw Error(MGMC_ERR_ARG, "cannot coarsen lattice [no interior vertices]");
    nc[d] = n[d] / 2;
  }
  return make(dim, nc);
}

void Geom::euclid(int64_t ell, int* idx) const {
  idx[0] = (int)(ell % (n[0] - 1)) + 1;
  ell /= (n[0] - 1);
  idx[1] = (int)(ell % (n[1] - 1)) + 1;
  ell /= (n[1] - 1);
  idx[2] = dim == 3 ? (int)ell + 1 : 0;
}

int64_t Geom::padded(int64_t ell) const {
  int e[3];
  euclid(ell, e);
  return ((int64_t)e[2] * py + e[1]) * px + e[0];
}

static inline int sidx(int dx, int dy, int dz) { return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); }

// RAP of a constant stencil: s_c(D) = sum_{sigma,tau} p(sigma) p(tau) s(2D + tau - sigma)
void galerkin_stencil(int dim, const double* s, double* c) {
  const double w1[3] = {0.5, 1.0, 0.5};
  int zr = dim == 3 ? 1 : 0;
  for (int i = 0; i < 27; ++i) c[i] = 0.0;
  for (int Dz = -zr; Dz <= zr; ++Dz)
    for (int Dy = -1; Dy <= 1; ++Dy)
      for (int Dx = -1; Dx <= 1; ++Dx) {
        double acc = 0.0;
        for (int sz = -zr; sz <= zr; ++sz)
          for (int sy = -1; sy <= 1; ++sy)
            for (int sx = -1; sx <= 1; ++sx)
              for (int tz = -zr; tz <= zr; ++tz)
                for (int ty = -1; ty <= 1; ++ty)
                  for (int tx = -1; tx <= 1; ++tx) {
                    int ox = 2 * Dx + tx - sx, oy = 2 * Dy + ty - sy, oz = 2 * Dz + tz - sz;
                    if (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1) continue;
                    double p = w1[sx + 1] * w1[sy + 1] * w1[tx + 1] * w1[ty + 1];
                    if (dim == 3) p *= w1[sz + 1] * w1[tz + 1];
                    acc += p * s[sidx(ox, oy, oz)];
                  }
        c[sidx(Dx, Dy, Dz)] = acc;
      }
}

static void classify(HostLevel& L) {
  if (L.variable) {  // star-ness is decided when the planes are built
    L.ncolour = L.star ? 2 : (1 << L.g.dim);
    return;
  }
  L.star = true;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int nz = (dx != 0) + (dy != 0) + (dz != 0);
        if (nz > 1 && L.st[sidx(dx, dy, dz)] != 0.0) L.star = false;
      }
  L.ncolour = L.star ? 2 : (1 << L.g.dim);
  L.diag = L.st[sidx(0, 0, 0)];
  if (!(L.diag > 0.0)) throw Error(MGMC_ERR_UNSUPPORTED, "stencil diagonal must be positive");
}

// ---- variable coefficients ----

// Interior vertices of a level in reference order, with padded index.
template <typename F>
static void for_interior(const Geom& g, F&& f) {
  const int nk = g.dim == 3 ? g.n[2] - 1 : 1;
  for (int k = 1; k <= nk; ++k)
    for (int j = 1; j <= g.n[1] - 1; ++j)
      for (int i = 1; i <= g.n[0] - 1; ++i) {
        const int kk = g.dim == 3 ? k : 0;
        f(i, j, kk, ((int64_t)kk * g.py + j) * g.px + i);
      }
}

static void check_positive_diagonal(const HostLevel& L) {
  bool ok = true;
  for_interior(L.g, [&](int, int, int, int64_t p) { ok = ok && L.coef(p, 13) > 0.0; });
  if (!ok) throw Error(MGMC_ERR_UNSUPPORTED, "stencil diagonal must be positive");
}

// -div(a grad) + kappa^2 by finite differences on the vertices, scaled by the
// cell volume as ShiftedLaplaceFDOperator (src/linear_operator/shiftedlaplace_fd_operator.cc:33-56):
// the coupling of a vertex to its neighbour along axis d is -h^d a_e / h_d^2
// with a_e the mean of a over the 2^(d-1) cells that share the edge, and the
// diagonal h^d (kappa^2 + sum over the 2 d edges of a_e / h_d^2), edges to
// boundary vertices included (Dirichlet elimination). For a = 1 and constant
// kappa^2 this is the reference's operator, diagonal accumulated in the same order.
static void fd_variable(HostLevel& L, double kappa_sq, const double* a_cell, const double* kappa_field) {
  const Geom& g = L.g;
  L.variable = true;
  L.star = true;
  L.vst.assign((size_t)kStarPlanes * g.vol, 0.0);
  double hinv2[3] = {0, 0, 0}, cv = 1.0;
  for (int d = 0; d < g.dim; ++d) {
    const double h = 1. / double(g.n[d]);
    hinv2[d] = 1. / (h * h);
    cv *= h;
  }
  auto a_at = [&](int ci, int cj, int ck) -> double {  // cell (ci, cj, ck), 0-based
    if (!a_cell) return 1.0;
    return a_cell[((int64_t)(g.dim == 3 ? ck : 0) * g.n[1] + cj) * g.n[0] + ci];
  };
  // mean of a over the cells sharing the edge from vertex (i, j, k) to its + neighbour along d
  auto edge = [&](int d, int i, int j, int k) -> double {
    const int e[3] = {i, j, k};
    int lo[3], hi[3];
    for (int q = 0; q < 3; ++q) {
      if (q == d) lo[q] = hi[q] = e[q];
      else if (q < g.dim) {
        lo[q] = e[q] - 1;
        hi[q] = e[q];
      } else lo[q] = hi[q] = 0;
    }
    double sum = 0.0;
    int cnt = 0;
    for (int ck = lo[2]; ck <= hi[2]; ++ck)
      for (int cj = lo[1]; cj <= hi[1]; ++cj)
        for (int ci = lo[0]; ci <= hi[0]; ++ci) {
          sum += a_at(ci, cj, ck);
          ++cnt;
        }
    return sum / cnt;
  };
  int64_t ref = 0;
  for_interior(g, [&](int i, int j, int k, int64_t p) {
    const double kap = kappa_field ? kappa_field[ref] : kappa_sq;
    ++ref;
    double diagonal = cv * kap;
    const int e[3] = {i, j, k};
    for (int d = 0; d < g.dim; ++d) {
      int m[3] = {e[0], e[1], e[2]};
      m[d] -= 1;
      const double am = edge(d, m[0], m[1], m[2]), ap = edge(d, e[0], e[1], e[2]);
      const int qm = 2 - d, qp = 4 + d;  // planes -x,-y,-z = 2,1,0; +x,+y,+z = 4,5,6
      L.vst[(size_t)qm * g.vol + p] = -cv * hinv2[d] * am;
      L.vst[(size_t)qp * g.vol + p] = -cv * hinv2[d] * ap;
      diagonal += cv * hinv2[d] * (am + ap);
    }
    L.vst[(size_t)3 * g.vol + p] = diagonal;
  });
  // neighbours outside the interior are eliminated: their couplings multiply
  // the zero boundary values, so they are kept (constant-stencil convention)
}

// A caller-supplied CSR matrix that is not one constant stencil: per-vertex
// stencils, checked to couple nearest neighbours only and to be symmetric.
static void stencils_from_csr(HostLevel& L, const int64_t* rp, const int64_t* ci, const double* v) {
  const Geom& g = L.g;
  std::vector<double> full((size_t)27 * g.vol, 0.0);
  bool star = true;
  double scale = 0.0;
  for (int64_t r = 0; r < g.N; ++r) {
    int a[3];
    g.euclid(r, a);
    const int64_t p = ((int64_t)a[2] * g.py + a[1]) * g.px + a[0];
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      int b[3];
      g.euclid(ci[k], b);
      const int o[3] = {b[0] - a[0], b[1] - a[1], g.dim == 3 ? b[2] - a[2] : 0};
      for (int d = 0; d < 3; ++d)
        if (o[d] < -1 || o[d] > 1) throw Error(MGMC_ERR_UNSUPPORTED, "A couples vertices that are not neighbours");
      full[(size_t)sidx(o[0], o[1], o[2]) * g.vol + p] += v[k];
      if (v[k] != 0.0 && (o[0] != 0) + (o[1] != 0) + (o[2] != 0) > 1) star = false;
      scale = std::max(scale, std::fabs(v[k]));
    }
  }
  const double tol = 1e-11 * scale;
  const int64_t pl = (int64_t)g.px * g.py;
  for_interior(g, [&](int, int, int, int64_t p) {
    for (int e = 0; e < 27; ++e) {
      const int dz = e / 9 - 1, dy = (e / 3) % 3 - 1, dx = e % 3 - 1;
      if (g.dim == 2 && dz != 0) continue;
      const int64_t q = p + dz * pl + dy * g.px + dx;
      if (q < 0 || q >= g.vol) continue;
      const double back = full[(size_t)(26 - e) * g.vol + q];  // entry (q, p)
      if (back != 0.0 || full[(size_t)e * g.vol + p] != 0.0)
        if (std::fabs(full[(size_t)e * g.vol + p] - back) > tol) throw Error(MGMC_ERR_UNSUPPORTED, "A is not symmetric");
    }
  });
  L.variable = true;
  L.star = star;
  if (star) {
    L.vst.assign((size_t)kStarPlanes * g.vol, 0.0);
    for (int q = 0; q < kStarPlanes; ++q)
      std::copy(full.begin() + (size_t)kStarIndex[q] * g.vol, full.begin() + (size_t)(kStarIndex[q] + 1) * g.vol,
                L.vst.begin() + (size_t)q * g.vol);
  } else {
    L.vst.swap(full);
  }
}

// Galerkin coarse stencils R A R^T of a variable level (R = P^T, P the tensor
// product of {1/2, 1, 1/2}; src/linear_operator/linear_operator.cc:10-23):
// for coarse vertex I, fine vertex i = 2I + s (weight p(s)) and its coupling
// o to j = i + o, j interpolates to the coarse vertices J within one step
// (weight p(j - 2J)), so A_c(I, J) = sum_s sum_o sum_J p(s) A(i, o) p(j - 2J).
static void galerkin_variable(const HostLevel& F, HostLevel& Cc) {
  const Geom& gf = F.g;
  const Geom& gc = Cc.g;
  Cc.variable = true;
  Cc.star = false;
  Cc.vst.assign((size_t)27 * gc.vol, 0.0);
  const int zr = gf.dim == 3 ? 1 : 0;
  const double w1[3] = {0.5, 1.0, 0.5};
  for_interior(gc, [&](int I, int J, int K, int64_t pc) {
    const int C3[3] = {I, J, K};
    for (int sz = -zr; sz <= zr; ++sz)
      for (int sy = -1; sy <= 1; ++sy)
        for (int sx = -1; sx <= 1; ++sx) {
          const int f3[3] = {2 * I + sx, 2 * J + sy, gf.dim == 3 ? 2 * K + sz : 0};
          const double ps = w1[sx + 1] * w1[sy + 1] * (gf.dim == 3 ? w1[sz + 1] : 1.0);
          const int64_t pf = ((int64_t)f3[2] * gf.py + f3[1]) * gf.px + f3[0];
          for (int oz = -zr; oz <= zr; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
              for (int ox = -1; ox <= 1; ++ox) {
                const double a = F.coef(pf, sidx(ox, oy, oz));
                if (a == 0.0) continue;
                const int j3[3] = {f3[0] + ox, f3[1] + oy, f3[2] + oz};
                // coarse vertices J that j interpolates to, per direction
                int Jl[3][2], nJ[3];
                double wJ[3][2];
                for (int d = 0; d < 3; ++d) {
                  if (d >= gf.dim) {
                    Jl[d][0] = 0;
                    wJ[d][0] = 1.0;
                    nJ[d] = 1;
                  } else if (j3[d] % 2 == 0) {
                    Jl[d][0] = j3[d] / 2;
                    wJ[d][0] = 1.0;
                    nJ[d] = 1;
                  } else {
                    Jl[d][0] = (j3[d] - 1) / 2;
                    Jl[d][1] = (j3[d] + 1) / 2;
                    wJ[d][0] = wJ[d][1] = 0.5;
                    nJ[d] = 2;
                  }
                }
                for (int a2 = 0; a2 < nJ[2]; ++a2)
                  for (int a1 = 0; a1 < nJ[1]; ++a1)
                    for (int a0 = 0; a0 < nJ[0]; ++a0) {
                      const int D[3] = {Jl[0][a0] - C3[0], Jl[1][a1] - C3[1], gf.dim == 3 ? Jl[2][a2] - C3[2] : 0};
                      const double w = wJ[0][a0] * wJ[1][a1] * wJ[2][a2];
                      Cc.vst[(size_t)sidx(D[0], D[1], D[2]) * gc.vol + pc] += ps * a * w;
                    }
              }
        }
  });
}

// Fine stencil from ShiftedLaplaceFDOperator with constant kappa^2
// (shiftedlaplace_fd_operator.cc:21-56; the diagonal accumulates in the same order).
static void fd_stencil(const Geom& g, double kappa_sq, double* st) {
  for (int i = 0; i < 27; ++i) st[i] = 0.0;
  double hinv2[3], cv = 1.0;
  for (int d = 0; d < g.dim; ++d) {
    double h = 1. / double(g.n[d]);
    hinv2[d] = 1. / (h * h);
    cv *= h;
  }
  double diagonal = cv * kappa_sq;
  for (int d = 0; d < g.dim; ++d) {
    int o[3] = {0, 0, 0};
    o[d] = -1;
    st[sidx(o[0], o[1], o[2])] = -cv * hinv2[d];
    o[d] = 1;
    st[sidx(o[0], o[1], o[2])] = -cv * hinv2[d];
    diagonal += 2. * cv * hinv2[d];
  }
  st[sidx(0, 0, 0)] = diagonal;
}

// Recognise a caller-supplied CSR matrix as a constant stencil with
// Dirichlet elimination; returns false when the rows differ (the matrix is
// then taken as per-vertex stencils). Couplings beyond nearest neighbours
// are unsupported.
static bool stencil_from_csr(const Geom& g, const int64_t* rp, const int64_t* ci, const double* v, double* st) {
  for (int i = 0; i < 27; ++i) st[i] = 0.0;
  int zr = g.dim == 3 ? 1 : 0;
  // first pass: a row whose whole 3^d neighbourhood is interior defines the stencil
  int64_t ref_row = -1;
  for (int64_t r = 0; r < g.N && ref_row < 0; ++r) {
    int e[3];
    g.euclid(r, e);
    bool deep = e[0] > 1 && e[0] < g.n[0] - 1 && e[1] > 1 && e[1] < g.n[1] - 1;
    if (g.dim == 3) deep = deep && e[2] > 1 && e[2] < g.n[2] - 1;
    if (deep) ref_row = r;
  }
  if (ref_row < 0) ref_row = 0;
  double scale = 0.0;
  for (int64_t r = 0; r < g.N; ++r) {
    int a[3];
    g.euclid(r, a);
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      int b[3];
      if (ci[k] < 0 || ci[k] >= g.N) throw Error(MGMC_ERR_ARG, "A column index out of range");
      g.euclid(ci[k], b);
      int o[3] = {b[0] - a[0], b[1] - a[1], g.dim == 3 ? b[2] - a[2] : 0};
      for (int d = 0; d < 3; ++d)
        if (o[d] < -1 || o[d] > 1) throw Error(MGMC_ERR_UNSUPPORTED, "A couples vertices that are not neighbours");
      if (r == ref_row) {
        st[sidx(o[0], o[1], o[2])] = v[k];
        scale = std::max(scale, std::fabs(v[k]));
      }
    }
  }
  (void)zr;
  // second pass: every row must be the stencil restricted to interior neighbours
  double tol = 1e-11 * scale;
  for (int64_t r = 0; r < g.N; ++r) {
    int a[3];
    g.euclid(r, a);
    double seen[27] = {0};
    bool hit[27] = {false};
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      int b[3];
      g.euclid(ci[k], b);
      int s = sidx(b[0] - a[0], b[1] - a[1], g.dim == 3 ? b[2] - a[2] : 0);
      seen[s] += v[k];
      hit[s] = true;
    }
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (g.dim == 2 && dz != 0) continue;
          int s = sidx(dx, dy, dz);
          int b[3] = {a[0] + dx, a[1] + dy, a[2] + dz};
          bool inside = b[0] >= 1 && b[0] <= g.n[0] - 1 && b[1] >= 1 && b[1] <= g.n[1] - 1;
          if (g.dim == 3) inside = inside && b[2] >= 1 && b[2] <= g.n[2] - 1;
          double want = inside ? st[s] : 0.0;
          if (std::fabs(seen[s] - want) > tol || (hit[s] && !inside)) return false;
        }
  }
  for (int s = 0; s < 27; ++s) {
    int dz = s / 9 - 1, dy = (s / 3) % 3 - 1, dx = s % 3 - 1;
    int t = sidx(-dx, -dy, -dz);
    if (std::fabs(st[s] - st[t]) > tol) throw Error(MGMC_ERR_UNSUPPORTED, "A is not symmetric");
  }
  return true;
}

static int colour_of(const HostLevel& L, int i, int j, int k) {
  if (L.star) return (i + j + k) & 1;
  return (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
}

// Gauss-Jordan inverse with partial pivoting (m x m row-major)
static std::vector<double> inverse(std::vector<double> A, int m) {
  std::vector<double> I(m * m, 0.0);
  for (int i = 0; i < m; ++i) I[i * m + i] = 1.0;
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(A[r * m + c]) > std::fabs(A[p * m + c])) p = r;
    if (A[p * m + c] == 0.0) throw Error(MGMC_ERR_NOT_SPD, "singular SMW capacitance matrix");
    if (p != c)
      for (int k = 0; k < m; ++k) {
        std::swap(A[p * m + k], A[c * m + k]);
        std::swap(I[p * m + k], I[c * m + k]);
      }
    double d = 1.0 / A[c * m + c];
    for (int k = 0; k < m; ++k) {
      A[c * m + k] *= d;
      I[c * m + k] *= d;
    }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      double f = A[r * m + c];
      if (f == 0.0) continue;
      for (int k = 0; k < m; ++k) {
        A[r * m + k] -= f * A[c * m + k];
        I[r * m + k] -= f * I[c * m + k];
      }
    }
  }
  return I;
}

// Column k of B_l: (padded index, value) sorted by index
using Column = std::vector<std::pair<int64_t, double>>;

static void set_columns(HostLevel& L, const std::vector<Column>& cols) {
  L.bptr.assign(1, 0);
  L.bidx.clear();
  L.bval.clear();
  std::map<int64_t, std::vector<std::pair<int, double>>> rows;
  for (size_t k = 0; k < cols.size(); ++k) {
    for (auto& e : cols[k]) {
      L.bidx.push_back(e.first);
      L.bval.push_back(e.second);
      rows[e.first].push_back({(int)k, e.second});
    }
    L.bptr.push_back((int32_t)L.bidx.size());
  }
  SparseRows& R = L.Brows;
  R = SparseRows();
  for (auto& kv : rows) {
    R.row.push_back(kv.first);
    for (auto& e : kv.second) {
      R.col.push_back(e.first);
      R.val.push_back(e.second);
    }
    R.ptr.push_back((int32_t)R.col.size());
  }
}

// B_{l+1} = R B_l (scatter form of src/intergrid/intergrid_operator.hh:74-88)
static std::vector<Column> restrict_columns(const Geom& f, const Geom& c, const std::vector<Column>& cols) {
  std::vector<Column> out(cols.size());
  for (size_t k = 0; k < cols.size(); ++k) {
    std::map<int64_t, double> acc;
    for (auto& e : cols[k]) {
      int64_t p = e.first;
      int i = (int)(p % f.px);
      int64_t q = p / f.px;
      int j = (int)(q % f.py);
      int kz = (int)(q / f.py);
      int I0[3], I1[3];
      double w0[3], w1[3];
      int cnt[3];
      int e3[3] = {i, j, kz};
      for (int d = 0; d < 3; ++d) {
        if (d >= f.dim) {
          I0[d] = 0; w0[d] = 1.0; cnt[d] = 1;
          continue;
        }
        if (e3[d] % 2 == 0) {
          I0[d] = e3[d] / 2; w0[d] = 1.0; cnt[d] = 1;
        } else {
          I0[d] = (e3[d] - 1) / 2; w0[d] = 0.5;
          I1[d] = (e3[d] + 1) / 2; w1[d] = 0.5; cnt[d] = 2;
        }
      }
      for (int a = 0; a < cnt[0]; ++a)
        for (int b = 0; b < cnt[1]; ++b)
          for (int g = 0; g < cnt[2]; ++g) {
            int ci = a ? I1[0] : I0[0], cj = b ? I1[1] : I0[1], ck = g ? I1[2] : I0[2];
            double w = (a ? w1[0] : w0[0]) * (b ? w1[1] : w0[1]) * (g ? w1[2] : w0[2]);
            if (ci < 1 || ci > c.n[0] - 1 || cj < 1 || cj > c.n[1] - 1) continue;
            if (c.dim == 3 && (ck < 1 || ck > c.n[2] - 1)) continue;
            acc[((int64_t)ck * c.py + cj) * c.px + ci] += w * e.second;
          }
    }
    for (auto& kv : acc) out[k].push_back(kv);
  }
  return out;
}

// W = T^{-1} B with T = D/omega + couplings to earlier colours (the
// triangular factor of SORSmoother for the multicolour visiting order).
static void compute_lowrank(HostLevel& L, const std::vector<Column>& cols, const std::vector<double>& sigma,
                            double omega) {
  const Geom& g = L.g;
  int m = (int)cols.size();
  int zr = g.dim == 3 ? 1 : 0;
  for (int dir = 0; dir < 2; ++dir) {
    std::map<int64_t, std::vector<std::pair<int, double>>> rows;
    for (int k = 0; k < m; ++k) {
      int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
      for (auto& e : cols[k]) {
        int64_t p = e.first;
        int c3[3] = {(int)(p % g.px), (int)((p / g.px) % g.py), (int)(p / ((int64_t)g.px * g.py))};
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], c3[d]);
          hi[d] = std::max(hi[d], c3[d]);
        }
      }
      if (cols[k].empty()) continue;
      int grow = L.ncolour;
      int b0[3], b1[3];
      for (int d = 0; d < 3; ++d) {
        int nmax = d < g.dim ? g.n[d] - 1 : 0;
        int nmin = d < g.dim ? 1 : 0;
        b0[d] = std::max(nmin, lo[d] - grow);
        b1[d] = std::min(nmax, hi[d] + grow);
      }
      int ex = b1[0] - b0[0] + 3, ey = b1[1] - b0[1] + 3, ez = b1[2] - b0[2] + 1 + 2 * zr;
      auto bi = [&](int i, int j, int kk) -> int64_t {
        return ((int64_t)(kk - b0[2] + zr) * ey + (j - b0[1] + 1)) * ex + (i - b0[0] + 1);
      };
      std::vector<double> bb((size_t)ex * ey * ez, 0.0), w((size_t)ex * ey * ez, 0.0);
      for (auto& e : cols[k]) {
        int64_t p = e.first;
        bb[bi((int)(p % g.px), (int)((p / g.px) % g.py), (int)(p / ((int64_t)g.px * g.py)))] += e.second;
      }
      for (int s = 0; s < L.ncolour; ++s) {
        int col = dir == 1 ? s : L.ncolour - 1 - s;
        for (int kk = b0[2]; kk <= b1[2]; ++kk)
          for (int j = b0[1]; j <= b1[1]; ++j)
            for (int i = b0[0]; i <= b1[0]; ++i) {
              if (colour_of(L, i, j, kk) != col) continue;
              double acc = bb[bi(i, j, kk)];
              const int64_t pv = ((int64_t)kk * g.py + j) * g.px + i;
              for (int dz = -zr; dz <= zr; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                  for (int dx = -1; dx <= 1; ++dx) {
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    double a = L.coef(pv, sidx(dx, dy, dz));
                    if (a != 0.0) acc -= a * w[bi(i + dx, j + dy, kk + dz)];
                  }
              w[bi(i, j, kk)] = acc / (L.coef(pv, 13) / omega);
            }
      }
      for (int kk = b0[2]; kk <= b1[2]; ++kk)
        for (int j = b0[1]; j <= b1[1]; ++j)
          for (int i = b0[0]; i <= b1[0]; ++i) {
            double v = w[bi(i, j, kk)];
            if (v != 0.0) rows[((int64_t)kk * g.py + j) * g.px + i].push_back({k, v});
          }
    }
    SparseRows& W = L.W[dir];
    W = SparseRows();
    std::unordered_map<int64_t, int> where;
    for (auto& kv : rows) {
      where[kv.first] = W.nrows();
      W.row.push_back(kv.first);
      for (auto& e : kv.second) {
        W.col.push_back(e.first);
        W.val.push_back(e.second);
      }
      W.ptr.push_back((int32_t)W.col.size());
    }
    // G = B^T W, S = Sigma + G
    std::vector<double> S((size_t)m * m, 0.0);
    for (int k = 0; k < m; ++k) {
      for (auto& e : cols[k]) {
        auto it = where.find(e.first);
        if (it == where.end()) continue;
        int r = it->second;
        for (int a = W.ptr[r]; a < W.ptr[r + 1]; ++a) S[(size_t)k * m + W.col[a]] += e.second * W.val[a];
      }
    }
    L.G[dir] = S;
    for (int k = 0; k < m; ++k) S[(size_t)k * m + k] += sigma[k];
    L.Sinv[dir] = inverse(S, m);
  }
}

// (Sigma + B^T T^{-1} B)^{-1} for the lexicographic visiting order, where T
// = D/omega + the couplings to the vertices visited before (earlier index
// forward, dir = 1; later index backward, dir = 0): the capacitance matrix of
// SORSmoother's SMW correction (src/smoother/sor_smoother.cc:17-37) in the
// reference's own order. Each column w = T^{-1} b_k is one forward (backward)
// substitution over the whole level, as the reference's triangular solve.
void lex_capacitance(const HostLevel& L, double omega, const std::vector<double>& sigma, std::vector<double> out[2]) {
  const Geom& g = L.g;
  const int m = L.m;
  const int zr = g.dim == 3 ? 1 : 0;
  const int64_t pl = (int64_t)g.px * g.py;
  // the box plus one plane, row and slot: when px = n_x the (+1, +1, +1)
  // neighbour of the last vertex is the slot after the box (zero, as the
  // padding after a chain's box is on the device)
  const size_t len = (size_t)g.vol + (size_t)pl + (size_t)g.px + 1;
  std::vector<double> b(len), w(len);
  for (int dir = 0; dir < 2; ++dir) {
    // stencil entries of the neighbours visited before a vertex, increasing index order
    std::vector<std::pair<int64_t, int>> before;
    for (int dz = -zr; dz <= zr; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const bool earlier = dz < 0 || (dz == 0 && (dy < 0 || (dy == 0 && dx < 0)));
          if ((dx || dy || dz) && earlier == (dir == 1)) before.push_back({dz * pl + dy * g.px + dx, sidx(dx, dy, dz)});
        }
    std::vector<double> S((size_t)m * m, 0.0);
    for (int l = 0; l < m; ++l) {
      std::fill(b.begin(), b.end(), 0.0);
      std::fill(w.begin(), w.end(), 0.0);
      for (int32_t e = L.bptr[l]; e < L.bptr[l + 1]; ++e) b[L.bidx[e]] += L.bval[e];
      const int64_t nvtx = g.N;
      for (int64_t q = 0; q < nvtx; ++q) {
        const int64_t v = dir == 1 ? q : nvtx - 1 - q;
        const int64_t p = g.padded(v);
        double acc = b[p];
        for (auto& o : before) {
          const double a = L.coef(p, o.second);
          if (a != 0.0) acc -= a * w[p + o.first];
        }
        w[p] = acc / (L.coef(p, 13) / omega);
      }
      for (int k = 0; k < m; ++k)
        for (int32_t e = L.bptr[k]; e < L.bptr[k + 1]; ++e) S[(size_t)k * m + l] += L.bval[e] * w[L.bidx[e]];
    }
    for (int k = 0; k < m; ++k) S[(size_t)k * m + k] += sigma[k];
    out[dir] = inverse(S, m);
  }
}

static void build_coarse(const HostLevel& L, const std::vector<Column>& cols, const std::vector<double>& sigma,
                         HostCoarse& C) {
  const Geom& g = L.g;
  int64_t Nc = g.N;
  if (Nc > kMaxCoarse) throw Error(MGMC_ERR_UNSUPPORTED, "coarsest level too large for the dense Cholesky sampler");
  C.Nc = Nc;
  C.pidx.resize(Nc);
  std::unordered_map<int64_t, int64_t> ref;
  for (int64_t r = 0; r < Nc; ++r) {
    C.pidx[r] = g.padded(r);
    ref[C.pidx[r]] = r;
  }
  std::vector<double> Q((size_t)Nc * Nc, 0.0);
  int zr = g.dim == 3 ? 1 : 0;
  for (int64_t r = 0; r < Nc; ++r) {
    int e[3];
    g.euclid(r, e);
    for (int dz = -zr; dz <= zr; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int b[3] = {e[0] + dx, e[1] + dy, e[2] + dz};
          if (b[0] < 1 || b[0] > g.n[0] - 1 || b[1] < 1 || b[1] > g.n[1] - 1) continue;
          if (g.dim == 3 && (b[2] < 1 || b[2] > g.n[2] - 1)) continue;
          int64_t p = ((int64_t)b[2] * g.py + b[1]) * g.px + b[0];
          Q[(size_t)r * Nc + ref[p]] += L.coef(C.pidx[r], sidx(dx, dy, dz));
        }
  }
  for (size_t k = 0; k < cols.size(); ++k)
    for (auto& a : cols[k])
      for (auto& b : cols[k]) Q[(size_t)ref[a.first] * Nc + ref[b.first]] += a.second * b.second / sigma[k];
  // Cholesky Q = L L^T
  std::vector<double>& A = Q;
  for (int64_t j = 0; j < Nc; ++j) {
    double s = A[j * Nc + j];
    for (int64_t k = 0; k < j; ++k) s -= A[j * Nc + k] * A[j * Nc + k];
    if (!(s > 0.0)) throw Error(MGMC_ERR_NOT_SPD, "coarse precision matrix is not positive definite");
    double d = std::sqrt(s);
    A[j * Nc + j] = d;
    for (int64_t i = j + 1; i < Nc; ++i) {
      double t = A[i * Nc + j];
      for (int64_t k = 0; k < j; ++k) t -= A[i * Nc + k] * A[j * Nc + k];
      A[i * Nc + j] = t / d;
    }
  }
  // Linv = L^{-1} (lower triangular), column by column
  std::vector<double> Linv((size_t)Nc * Nc, 0.0);
  for (int64_t c = 0; c < Nc; ++c)
    for (int64_t i = c; i < Nc; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int64_t k = c; k < i; ++k) s -= A[i * Nc + k] * Linv[k * Nc + c];
      Linv[i * Nc + c] = s / A[i * Nc + i];
    }
  C.LinvT.assign((size_t)Nc * Nc, 0.0);
  for (int64_t i = 0; i < Nc; ++i)
    for (int64_t j = 0; j < Nc; ++j) C.LinvT[i * Nc + j] = Linv[j * Nc + i];
  C.Qinv.assign((size_t)Nc * Nc, 0.0);
  for (int64_t i = 0; i < Nc; ++i)
    for (int64_t j = 0; j < Nc; ++j) {
      double s = 0.0;
      for (int64_t k = std::max(i, j); k < Nc; ++k) s += Linv[k * Nc + i] * Linv[k * Nc + j];
      C.Qinv[i * Nc + j] = s;
    }
}

HostHierarchy build_hierarchy(const HostProblem& p, int nlevel, double omega, bool coarse_cholesky) {
  if (nlevel < 1) throw Error(MGMC_ERR_ARG, "nlevel must be >= 1");
  if (!(omega > 0.0 && omega < 2.0)) throw Error(MGMC_ERR_ARG, "omega must lie in (0, 2)");
  if (p.m < 0) throw Error(MGMC_ERR_ARG, "m must be >= 0");
  HostHierarchy H;
  H.omega = omega;
  H.sigma.assign(p.sigma ? p.sigma : nullptr, p.sigma ? p.sigma + p.m : nullptr);
  for (int k = 0; k < p.m; ++k)
    if (!(H.sigma[k] > 0.0)) throw Error(MGMC_ERR_ARG, "measurement variances must be positive");
  Geom g = Geom::make(p.dim, p.n);
  HostLevel L0;
  L0.g = g;
  if (p.A_rowptr) {
    if (p.diffusion || p.kappa_sq_field) throw Error(MGMC_ERR_ARG, "give A either as CSR or by its coefficient fields");
    if (!stencil_from_csr(g, p.A_rowptr, p.A_col, p.A_val, L0.st))
      stencils_from_csr(L0, p.A_rowptr, p.A_col, p.A_val);
  } else if (p.diffusion || p.kappa_sq_field) {
    if (!p.kappa_sq_field && !(p.kappa_sq >= 0.0)) throw Error(MGMC_ERR_ARG, "kappa^2 must be >= 0");
    const int64_t ncell = (int64_t)g.n[0] * g.n[1] * (g.dim == 3 ? g.n[2] : 1);
    for (int64_t c = 0; p.diffusion && c < ncell; ++c)
      if (!(p.diffusion[c] > 0.0)) throw Error(MGMC_ERR_ARG, "the diffusion coefficient must be positive");
    for (int64_t v = 0; p.kappa_sq_field && v < g.N; ++v)
      if (!(p.kappa_sq_field[v] >= 0.0)) throw Error(MGMC_ERR_ARG, "kappa^2 must be >= 0");
    fd_variable(L0, p.kappa_sq, p.diffusion, p.kappa_sq_field);
  } else {
    if (!(p.kappa_sq >= 0.0)) throw Error(MGMC_ERR_ARG, "kappa^2 must be >= 0");
    fd_stencil(g, p.kappa_sq, L0.st);
  }
  classify(L0);
  if (L0.variable) check_positive_diagonal(L0);
  std::vector<Column> cols(p.m);
  if (p.m > 0) {
    if (!p.B_rowptr) throw Error(MGMC_ERR_ARG, "m > 0 needs B");
    for (int64_t r = 0; r < g.N; ++r)
      for (int64_t k = p.B_rowptr[r]; k < p.B_rowptr[r + 1]; ++k) {
        if (p.B_col[k] < 0 || p.B_col[k] >= p.m) throw Error(MGMC_ERR_ARG, "B column index out of range");
        cols[p.B_col[k]].push_back({g.padded(r), p.B_val[k]});
      }
    for (auto& c : cols) std::sort(c.begin(), c.end());
  }
  H.levels.push_back(L0);
  for (int l = 1; l < nlevel; ++l) {
    HostLevel& F = H.levels.back();
    HostLevel C;
    C.g = F.g.coarse();
    if (F.variable) galerkin_variable(F, C);
    else galerkin_stencil(C.g.dim, F.st, C.st);
    classify(C);
    if (C.variable) check_positive_diagonal(C);
    H.levels.push_back(C);
  }
  for (int l = 0; l < nlevel; ++l) {
    HostLevel& L = H.levels[l];
    L.m = p.m;
    if (l > 0 && p.m > 0) cols = restrict_columns(H.levels[l - 1].g, L.g, cols);
    if (p.m > 0) {
      set_columns(L, cols);
      compute_lowrank(L, cols, H.sigma, omega);
    }
    // The dense coarse factor serves the Cholesky sampler and the
    // preconditioner's coarse solve (always Cholesky, as in
    // multigrid_preconditioner.cc:41-45). With an SSOR coarse sampler it is
    // built only when the coarsest level is small enough.
    if (l == nlevel - 1 && (coarse_cholesky || L.g.N <= kMaxCoarseKernel)) build_coarse(L, cols, H.sigma, H.coarse);
  }
  return H;
}

}  // namespace mgmc
