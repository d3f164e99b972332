// Host-side operator setup that C/C++ callers of the library need without
// Python: the ball-average measurement vector of MeasuredOperator and the
// constant 3^d stencil of the Q1 finite-element prior. FP64, setup only.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "setup.h"

namespace mgmc {

// Volume of the d-ball (src/linear_operator/measured_operator.cc:52-66).
static double ball_volume(double r, int dim) {
  if (dim == 0) return 1.0;
  if (dim == 1) return 2.0 * r;
  return 2.0 * M_PI / dim * r * r * ball_volume(r, dim - 2);
}

// Reference vertex index of interior vertex (1-based Euclidean indices).
static int64_t vertex_linear(int dim, const int* n, const int* e) {
  int64_t ell = 0, stride = 1;
  for (int d = 0; d < dim; ++d) {
    ell += (int64_t)(e[d] - 1) * stride;
    stride *= n[d] - 1;
  }
  return ell;
}

// Integral of exp(-|u|^2 / (2 w^2)) over the unit d-ball (radial
// integration). 2D: 2 pi w^2 (1 - e^{-1/(2w^2)}), with expm1. 3D: 4 pi
// int_0^1 u^2 e^{-u^2/(2w^2)} du, which is w^3 (sqrt(pi/2) erf(1/(sqrt(2) w))
// - e^{-1/(2w^2)} / w) (by parts) for w < 1 and the series
// sum_k (-1)^k / ((2w^2)^k k! (2k + 3)) for w >= 1, where the closed form
// cancels.
double gaussian_ball_integral(double w, int dim) {
  const double a = 0.5 / (w * w);
  if (dim == 2) return -2.0 * M_PI * w * w * std::expm1(-a);
  if (w < 1.0)
    return 4.0 * M_PI * w * w * w * (std::sqrt(0.5 * M_PI) * std::erf(1.0 / (std::sqrt(2.0) * w)) - std::exp(-a) / w);
  double sum = 0.0, term = 1.0;  // term = (-a)^k / k!
  for (int k = 0; k < 40; ++k) {
    sum += term / (2 * k + 3);
    term *= -a / (k + 1);
  }
  return 4.0 * M_PI * sum;
}

// w_j = int f(x) phi_j(x) dx for a weight f of integral one on the ball
// (x0, radius), 2^d-point Gauss-Legendre per cell; the nearest vertex for
// radius < 1e-12 (src/linear_operator/measured_operator.cc:69-170). Cells are
// visited in increasing linear order inside a bounding box of the ball, a
// cell counts when one of its corners lies inside the ball or it contains
// x0, and contributions accumulate cell by cell, corner by corner.
// profile 0: the reference's top-hat f = 1 / V_d(R) (f_meas = 1,
// measured_operator.hh:65); profile 1: f = exp(-xi^2 / (2 width^2)) / (R^d c_d)
// with xi = |x - x0| / R < 1 and c_d = gaussian_ball_integral(width, d), a
// Gaussian-weighted ball average (no counterpart in the reference).
void measurement_vector(int dim, const int* n, const double* x0, double radius, std::vector<int64_t>& idx,
                        std::vector<double>& val, int profile, double width) {
  if (profile != 0 && profile != 1) throw Error(1, "measurement vector: profile must be 0 (top-hat) or 1 (Gaussian)");
  if (profile == 1 && !(width > 0.0)) throw Error(1, "measurement vector: the Gaussian width must be positive");
  idx.clear();
  val.clear();
  if (dim < 2 || dim > 3) throw Error(1, "measurement vector: dim must be 2 or 3");
  double h[3], cv = 1.0;
  for (int d = 0; d < dim; ++d) {
    h[d] = 1.0 / n[d];
    cv *= h[d];
  }
  if (radius < 1e-12) {
    double best = INFINITY;
    int64_t arg = -1;
    int e[3] = {1, 1, 1};
    const int n2 = dim == 3 ? n[2] - 1 : 1;
    for (int k = 1; k <= n2; ++k)
      for (int j = 1; j <= n[1] - 1; ++j)
        for (int i = 1; i <= n[0] - 1; ++i) {
          e[0] = i;
          e[1] = j;
          e[2] = k;
          double s = 0.0;
          for (int d = 0; d < dim; ++d) s += (e[d] * h[d] - x0[d]) * (e[d] * h[d] - x0[d]);
          const double dist = std::sqrt(s);
          if (dist < best) {
            best = dist;
            arg = vertex_linear(dim, n, e);
          }
        }
    idx.push_back(arg);
    val.push_back(1.0);
    return;
  }
  const double norm =
      profile == 0 ? 1.0 / ball_volume(radius, dim) : 1.0 / (std::pow(radius, dim) * gaussian_ball_integral(width, dim));
  const int nq = 1 << dim;
  const double g1[2] = {0.5 * (-1.0 / std::sqrt(3.0) + 1.0), 0.5 * (1.0 / std::sqrt(3.0) + 1.0)};
  const double qw = std::pow(0.5, dim);
  // corner / quadrature-point bits: entry c, component d = bit (dim-1-d) of c
  auto bit = [dim](int c, int d) { return (c >> (dim - 1 - d)) & 1; };
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    lo[d] = std::max(0, (int)std::floor((x0[d] - radius) / h[d]) - 1);
    hi[d] = std::min(n[d] - 1, (int)std::ceil((x0[d] + radius) / h[d]) + 1);
  }
  std::map<int64_t, double> acc;
  const int klo = dim == 3 ? lo[2] : 0, khi = dim == 3 ? hi[2] : 0;
  for (int ck = klo; ck <= khi; ++ck)
    for (int cj = lo[1]; cj <= hi[1]; ++cj)
      for (int ci = lo[0]; ci <= hi[0]; ++ci) {
        const int cell[3] = {ci, cj, ck};
        bool overlap = false;
        bool contains = true;
        for (int a = 0; a < nq; ++a) {
          double s = 0.0;
          for (int d = 0; d < dim; ++d) {
            const double c = h[d] * (cell[d] + bit(a, d));
            s += (c - x0[d]) * (c - x0[d]);
          }
          if (std::sqrt(s) < radius) overlap = true;
        }
        for (int d = 0; d < dim; ++d)
          contains = contains && h[d] * cell[d] <= x0[d] && x0[d] <= h[d] * (cell[d] + 1);
        if (!(overlap || contains)) continue;
        bool inside[8];
        double fq[8];
        for (int q = 0; q < nq; ++q) {
          double s = 0.0;
          for (int d = 0; d < dim; ++d) {
            const double xq = h[d] * (g1[bit(q, d)] + cell[d]);
            s += (xq - x0[d]) * (xq - x0[d]);
          }
          const double xi = std::sqrt(s) / radius;
          inside[q] = xi < 1.0;
          fq[q] = profile == 0 ? 1.0 : std::exp(-0.5 * xi * xi / (width * width));
        }
        for (int a = 0; a < nq; ++a) {
          int v[3] = {1, 1, 1};
          bool interior = true;
          for (int d = 0; d < dim; ++d) {
            v[d] = cell[d] + bit(a, d);
            interior = interior && v[d] >= 1 && v[d] <= n[d] - 1;
          }
          double local = 0.0;
          for (int q = 0; q < nq; ++q) {
            double phihat = fq[q];
            for (int d = 0; d < dim; ++d) {
              const double xq = g1[bit(q, d)];
              phihat *= bit(a, d) == 0 ? 1.0 - xq : xq;
            }
            local = local + (inside[q] ? phihat * qw * cv * norm : 0.0);
          }
          if (interior) acc[vertex_linear(dim, n, v)] += local;
        }
      }
  for (const auto& kv : acc) {
    idx.push_back(kv.first);
    val.push_back(kv.second);
  }
}

// 3^d stencil of Q1 finite elements for -Delta + kappa^2 on the lattice
// (src/linear_operator/shiftedlaplace_fem_operator.cc:71-121): element
// matrices by 2^d-point Gauss-Legendre quadrature, summed over the 2^d cells
// around a vertex. Output index (dz+1)*9 + (dy+1)*3 + (dx+1) (3D) or
// (dy+1)*3 + (dx+1) (2D).
void fem_stencil(int dim, const int* n, double kappa_sq, double* st) {
  if (dim < 2 || dim > 3) throw Error(1, "FEM stencil: dim must be 2 or 3");
  const int nb = 1 << dim;
  double h[3], hinv2[3], cv = 1.0;
  for (int d = 0; d < dim; ++d) {
    h[d] = 1.0 / n[d];
    hinv2[d] = 1.0 / (h[d] * h[d]);
    cv *= h[d];
  }
  const double g1[2] = {0.5 * (-1.0 / std::sqrt(3.0) + 1.0), 0.5 * (1.0 / std::sqrt(3.0) + 1.0)};
  const double qw = std::pow(0.5, dim);
  auto bit = [dim](int c, int d) { return (c >> (dim - 1 - d)) & 1; };
  auto phi = [&](int a, int q) {
    double v = 1.0;
    for (int d = 0; d < dim; ++d) v *= bit(a, d) ? g1[bit(q, d)] : 1.0 - g1[bit(q, d)];
    return v;
  };
  auto dphi = [&](int a, int q, int j) {
    double v = bit(a, j) ? 1.0 : -1.0;
    for (int d = 0; d < dim; ++d)
      if (d != j) v *= bit(a, d) ? g1[bit(q, d)] : 1.0 - g1[bit(q, d)];
    return v;
  };
  std::vector<double> M((size_t)nb * nb, 0.0);
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b) {
      double v = 0.0;
      for (int q = 0; q < nb; ++q) {
        double gg = 0.0;
        for (int j = 0; j < dim; ++j) gg += dphi(a, q, j) * hinv2[j] * dphi(b, q, j);
        v += (kappa_sq * phi(a, q) * phi(b, q) + gg) * qw;
      }
      M[(size_t)a * nb + b] = v * cv;
    }
  const int ns = dim == 3 ? 27 : 9;
  for (int e = 0; e < ns; ++e) st[e] = 0.0;
  // vertex = corner a of a cell; its neighbour at offset s is corner b = a + s
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b) {
      int s[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) s[d] = bit(b, d) - bit(a, d);
      const int e = dim == 3 ? (s[2] + 1) * 9 + (s[1] + 1) * 3 + (s[0] + 1) : (s[1] + 1) * 3 + (s[0] + 1);
      st[e] += M[(size_t)a * nb + b];
    }
}

}  // namespace mgmc
