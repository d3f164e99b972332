// FP64 CPU restatement of the MultigridMC multigrid Monte Carlo (MGMC) path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product path (multigridmc_amd) never links, loads or calls it.
//
// Every function names the reference file:line whose behaviour it restates
// (paths relative to the reference repository root).  The restatement keeps
// the reference's arithmetic: lexicographic SOR, dense Sherman-Morrison-
// Woodbury (SMW) correction matrices, Galerkin coarsening by sparse triple
// products, L'Ecuyer CLCG random numbers with Marsaglia's polar method.
// One extension is added so that the GPU kernels have an exact CPU twin:
// a "multicolour" ordering option (red-black for star stencils, 2^d colours
// for box stencils) that changes only the order in which SOR visits the
// unknowns (and hence the triangular factor used in the SMW correction).
//
// The C ABI at the end of this file is what the Python tests bind via ctypes.
// Each function is pinned by a check in tests/test_oracle_*.py against the
// reference's known-answer tests, closed forms or brute force; none is left
// unpinned.

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <random>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Lattice: src/lattice/lattice{1d,2d,3d}.hh
// nx = number of cells; unknowns on interior vertices 1..n-1 (x fastest).
// ---------------------------------------------------------------------------
struct Lattice {
  int dim = 0;
  int n[3] = {1, 1, 1};

  long Nvertex() const {
    long N = 1;
    for (int d = 0; d < dim; ++d) N *= (n[d] - 1);
    return N;
  }
  long Ncell() const {
    long N = 1;
    for (int d = 0; d < dim; ++d) N *= n[d];
    return N;
  }
  double cell_volume() const {  // lattice.hh:34-43
    double v = 1.0;
    for (int d = 0; d < dim; ++d) v /= n[d];
    return v;
  }
  // vertexidx_linear2euclidean (lattice3d.hh:118-126)
  void v2e(long ell, int* idx) const {
    for (int d = 0; d < dim; ++d) {
      idx[d] = int(ell % (n[d] - 1)) + 1;
      ell /= (n[d] - 1);
    }
  }
  // vertexidx_euclidean2linear (lattice3d.hh:132-141); -1 if not interior
  long e2v(const int* idx) const {
    long ell = 0, stride = 1;
    for (int d = 0; d < dim; ++d) {
      if (idx[d] <= 0 || idx[d] >= n[d]) return -1;
      ell += long(idx[d] - 1) * stride;
      stride *= (n[d] - 1);
    }
    return ell;
  }
  // cellidx_linear2euclidean (lattice3d.hh:90-97)
  void c2e(long ell, int* idx) const {
    for (int d = 0; d < dim; ++d) {
      idx[d] = int(ell % n[d]);
      ell /= n[d];
    }
  }
  // vertex_coordinates (lattice3d.hh:237-245)
  void coords(long ell, double* x) const {
    int idx[3];
    v2e(ell, idx);
    for (int d = 0; d < dim; ++d) x[d] = idx[d] * (1.0 / double(n[d]));
  }
  // get_coarse_lattice (lattice3d.hh:248-263): errors instead of exit(-1)
  Lattice coarse() const {
    Lattice c;
    c.dim = dim;
    for (int d = 0; d < dim; ++d) {
      if (n[d] % 2 != 0) throw std::runtime_error("cannot coarsen lattice [one of the extents is odd]");
      if (!(n[d] / 2 > 1)) throw std::runtime_error("cannot coarsen lattice [no interior vertices]");
      c.n[d] = n[d] / 2;
    }
    return c;
  }
};

// ---------------------------------------------------------------------------
// Minimal CSR sparse matrix (Eigen::SparseMatrix stand-in, sorted columns)
// ---------------------------------------------------------------------------
struct CSR {
  long nrow = 0, ncol = 0;
  std::vector<long> rp{0};
  std::vector<long> ci;
  std::vector<double> v;

  struct Trip { long r, c; double v; };
  static CSR from_triplets(long nrow, long ncol, std::vector<Trip>& t) {
    // duplicates are summed, as Eigen's setFromTriplets does
    std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
      return a.r < b.r || (a.r == b.r && a.c < b.c);
    });
    CSR M;
    M.nrow = nrow;
    M.ncol = ncol;
    M.rp.assign(nrow + 1, 0);
    for (size_t k = 0; k < t.size();) {
      size_t e = k;
      double s = 0.0;
      while (e < t.size() && t[e].r == t[k].r && t[e].c == t[k].c) s += t[e++].v;
      M.ci.push_back(t[k].c);
      M.v.push_back(s);
      M.rp[t[k].r + 1]++;
      k = e;
    }
    for (long r = 0; r < nrow; ++r) M.rp[r + 1] += M.rp[r];
    return M;
  }
  double diag(long r) const {
    for (long k = rp[r]; k < rp[r + 1]; ++k)
      if (ci[k] == r) return v[k];
    return 0.0;
  }
  void matvec(const double* x, double* y) const {
    for (long r = 0; r < nrow; ++r) {
      double s = 0.0;
      for (long k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
      y[r] = s;
    }
  }
  // y = M^T x
  void matvec_T(const double* x, double* y) const {
    std::fill(y, y + ncol, 0.0);
    for (long r = 0; r < nrow; ++r)
      for (long k = rp[r]; k < rp[r + 1]; ++k) y[ci[k]] += v[k] * x[r];
  }
  CSR transpose() const {
    std::vector<Trip> t;
    t.reserve(v.size());
    for (long r = 0; r < nrow; ++r)
      for (long k = rp[r]; k < rp[r + 1]; ++k) t.push_back({ci[k], r, v[k]});
    return from_triplets(ncol, nrow, t);
  }
  // Gustavson product C = this * B
  CSR mul(const CSR& B) const {
    CSR C;
    C.nrow = nrow;
    C.ncol = B.ncol;
    C.rp.assign(nrow + 1, 0);
    std::vector<double> acc(B.ncol, 0.0);
    std::vector<char> used(B.ncol, 0);
    std::vector<long> cols;
    for (long r = 0; r < nrow; ++r) {
      cols.clear();
      for (long k = rp[r]; k < rp[r + 1]; ++k) {
        long j = ci[k];
        double a = v[k];
        for (long l = B.rp[j]; l < B.rp[j + 1]; ++l) {
          long c = B.ci[l];
          if (!used[c]) { used[c] = 1; cols.push_back(c); }
          acc[c] += a * B.v[l];
        }
      }
      std::sort(cols.begin(), cols.end());
      for (long c : cols) {
        C.ci.push_back(c);
        C.v.push_back(acc[c]);
        acc[c] = 0.0;
        used[c] = 0;
      }
      C.rp[r + 1] = (long)C.ci.size();
    }
    return C;
  }
};

// ---------------------------------------------------------------------------
// Dense helpers (Eigen dense LLT / inverse stand-ins)
// ---------------------------------------------------------------------------
// Cholesky A = L L^T in place (row-major n x n, lower part); false if not SPD
static bool cholesky(std::vector<double>& A, long n) {
  for (long j = 0; j < n; ++j) {
    double s = A[j * n + j];
    for (long k = 0; k < j; ++k) s -= A[j * n + k] * A[j * n + k];
    if (!(s > 0.0)) return false;
    double d = std::sqrt(s);
    A[j * n + j] = d;
    for (long i = j + 1; i < n; ++i) {
      double t = A[i * n + j];
      for (long k = 0; k < j; ++k) t -= A[i * n + k] * A[j * n + k];
      A[i * n + j] = t / d;
    }
    for (long i = 0; i < j; ++i) A[i * n + j] = 0.0;
  }
  return true;
}
static void solve_L(const std::vector<double>& L, long n, const double* b, double* x) {
  for (long i = 0; i < n; ++i) {
    double s = b[i];
    for (long k = 0; k < i; ++k) s -= L[i * n + k] * x[k];
    x[i] = s / L[i * n + i];
  }
}
static void solve_LT(const std::vector<double>& L, long n, const double* b, double* x) {
  for (long i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (long k = i + 1; k < n; ++k) s -= L[k * n + i] * x[k];
    x[i] = s / L[i * n + i];
  }
}
// General inverse by Gauss-Jordan with partial pivoting (row-major m x m)
static std::vector<double> inverse(std::vector<double> A, long m) {
  std::vector<double> I(m * m, 0.0);
  for (long i = 0; i < m; ++i) I[i * m + i] = 1.0;
  for (long c = 0; c < m; ++c) {
    long p = c;
    for (long r = c + 1; r < m; ++r)
      if (std::fabs(A[r * m + c]) > std::fabs(A[p * m + c])) p = r;
    if (A[p * m + c] == 0.0) throw std::runtime_error("singular matrix");
    if (p != c)
      for (long k = 0; k < m; ++k) {
        std::swap(A[p * m + k], A[c * m + k]);
        std::swap(I[p * m + k], I[c * m + k]);
      }
    double d = 1.0 / A[c * m + c];
    for (long k = 0; k < m; ++k) { A[c * m + k] *= d; I[c * m + k] *= d; }
    for (long r = 0; r < m; ++r) {
      if (r == c) continue;
      double f = A[r * m + c];
      if (f == 0.0) continue;
      for (long k = 0; k < m; ++k) { A[r * m + k] -= f * A[c * m + k]; I[r * m + k] -= f * I[c * m + k]; }
    }
  }
  return I;
}

// ---------------------------------------------------------------------------
// Random numbers: src/auxilliary/parallel_random.cc:9-77
// ---------------------------------------------------------------------------
struct CLCG {
  int32_t y1, y2;
  static constexpr int32_t a1 = 40014, m1 = 2147483563, a2 = 40692, m2 = 2147483399;
  static constexpr double m1_inv = 1. / 2147483563.;
  int64_t a1_multistep, a2_multistep;
  bool redraw = true;
  double z1 = 0, z2 = 0;
  // leapfrog construction for stream `id` of `nstreams` (parallel_random.cc:35-59)
  explicit CLCG(int nstreams = 1, int id = 0) {
    a1_multistep = a1;
    a2_multistep = a2;
    y1 = 1;
    y2 = 1;
    for (int j = 1; j < nstreams; ++j) {
      if (j <= id) {
        y1 = int32_t(((int64_t)a1 * y1) % m1);
        y2 = int32_t(((int64_t)a2 * y2) % m2);
      }
      a1_multistep = (a1_multistep * a1) % m1;
      a2_multistep = (a2_multistep * a2) % m2;
    }
  }
  int32_t draw_int() {  // parallel_random.cc:62-70
    int32_t x = y1 - y2;
    if (x < 1) x += m1 - 1;
    y1 = int32_t((a1_multistep * y1) % m1);
    y2 = int32_t((a2_multistep * y2) % m2);
    return x;
  }
  double draw_uniform() { return draw_int() * m1_inv; }  // :73-77
  double draw_normal() {                                // :9-32 (Marsaglia polar)
    if (redraw) {
      double u1, u2, s;
      do {
        u1 = 2. * draw_uniform() - 1.;
        u2 = 2. * draw_uniform() - 1.;
        s = u1 * u1 + u2 * u2;
      } while ((s == 0) || (s >= 1));
      double L = std::sqrt(-2 * std::log(s) / s);
      z1 = L * u1;
      z2 = L * u2;
      redraw = false;
      return z1;
    }
    redraw = true;
    return z2;
  }
};

// ---------------------------------------------------------------------------
// Operators: src/linear_operator/linear_operator.{hh,cc}
// Q = A + B Sigma^{-1} B^T
// ---------------------------------------------------------------------------
struct Op {
  Lattice lat;
  CSR A;
  CSR B;  // N x m
  std::vector<double> Sigma;
  int m = 0;

  long N() const { return A.nrow; }
  // LinearOperator::apply (linear_operator.hh:78-88)
  void apply(const double* x, double* y) const {
    A.matvec(x, y);
    if (m > 0) {
      std::vector<double> z(m);
      B.matvec_T(x, z.data());
      for (int k = 0; k < m; ++k) z[k] /= Sigma[k];
      for (long r = 0; r < B.nrow; ++r)
        for (long k = B.rp[r]; k < B.rp[r + 1]; ++k) y[r] += B.v[k] * z[B.ci[k]];
    }
  }
  std::vector<double> dense_precision() const {  // linear_operator.cc:26-34
    long n = N();
    std::vector<double> Q(n * n, 0.0);
    for (long r = 0; r < n; ++r)
      for (long k = A.rp[r]; k < A.rp[r + 1]; ++k) Q[r * n + A.ci[k]] += A.v[k];
    if (m > 0) {
      CSR BT = B.transpose();  // row k = column k of B
      for (int k = 0; k < m; ++k)
        for (long a = BT.rp[k]; a < BT.rp[k + 1]; ++a)
          for (long b = BT.rp[k]; b < BT.rp[k + 1]; ++b)
            Q[BT.ci[a] * n + BT.ci[b]] += BT.v[a] * BT.v[b] / Sigma[k];
    }
    return Q;
  }
};

// ShiftedLaplaceFDOperator ctor (shiftedlaplace_fd_operator.cc:9-57) with the
// constant correlation-length model (correlationlength_model.hh:45-66).
// kappa^2(x): constant model (correlationlength_model.hh:45-66) when
// Lambda_2 == 0, else the periodic model (correlationlength_model.hh:83-105)
struct KappaModel {
  double kappa_sq_const = 0;
  bool periodic = false;
  double L1 = 0, L2 = 0;
  double operator()(const double* x, int dim) const {
    if (!periodic) return kappa_sq_const;
    double Lam = L2;
    for (int d = 0; d < dim; ++d) Lam *= std::cos(M_PI * x[d]);
    Lam += L1;
    return 1. / (Lam * Lam);
  }
};

static Op make_fd_operator(const Lattice& lat, const KappaModel& kappa);
static Op make_fd_operator(const Lattice& lat, double kappa_sq) {
  KappaModel k;
  k.kappa_sq_const = kappa_sq;
  return make_fd_operator(lat, k);
}
static Op make_fd_operator(const Lattice& lat, const KappaModel& kappa) {
  Op op;
  op.lat = lat;
  int dim = lat.dim;
  long nrow = lat.Nvertex();
  double hinv2[3], cv = 1.0;
  for (int d = 0; d < dim; ++d) {
    double h = 1. / double(lat.n[d]);
    hinv2[d] = 1. / (h * h);
    cv *= h;
  }
  std::vector<CSR::Trip> t;
  t.reserve((1 + 2 * dim) * nrow);
  for (long ell = 0; ell < nrow; ++ell) {
    int idx[3];
    lat.v2e(ell, idx);
    double xc[3];
    lat.coords(ell, xc);
    double diagonal = cv * kappa(xc, dim);
    for (int d = 0; d < dim; ++d) {
      for (int j = 0; j < 2; ++j) {
        int s[3] = {idx[0], idx[1], idx[2]};
        s[d] += 2 * j - 1;
        long e = lat.e2v(s);
        if (e >= 0) t.push_back({ell, e, -cv * hinv2[d]});
      }
      diagonal += 2. * cv * hinv2[d];
    }
    t.push_back({ell, ell, diagonal});
  }
  op.A = CSR::from_triplets(nrow, nrow, t);
  return op;
}

// MeasuredOperator::V_sphere (measured_operator.cc:52-66)
static double V_sphere(double radius, int dim) {
  if (dim == 0) return 1.0;
  if (dim == 1) return 2. * radius;
  return 2. * M_PI / double(dim) * radius * radius * V_sphere(radius, dim - 2);
}

// MeasuredOperator::measurement_vector (measured_operator.cc:69-170), top-hat
// f_meas = 1 (measured_operator.hh:65), 2^d-point Gauss-Legendre per cell
// (quadrature.cc:11-56 with order 1). Returns a dense vector of length N.
static std::vector<double> measurement_vector(const Lattice& lat, const double* x0, double radius) {
  long N = lat.Nvertex();
  int dim = lat.dim;
  std::vector<double> r(N, 0.0);
  if (radius < 1.E-12) {
    double dmin = double(dim);
    long ellmin = 0;
    for (long ell = 0; ell < N; ++ell) {
      double x[3];
      lat.coords(ell, x);
      double s = 0;
      for (int d = 0; d < dim; ++d) s += (x[d] - x0[d]) * (x[d] - x0[d]);
      double dist = std::sqrt(s);
      if (dist < dmin) { dmin = dist; ellmin = ell; }
    }
    r[ellmin] = 1.0;
    return r;
  }
  double h[3];
  for (int d = 0; d < dim; ++d) h[d] = 1. / double(lat.n[d]);
  double cv = lat.cell_volume();
  double normalisation = 1. / V_sphere(radius, dim);
  // cartesian products: first component varies slowest (common.hh:29-52)
  int ncorner = 1 << dim;
  std::vector<std::array<int, 3>> corners(ncorner);
  std::vector<std::array<double, 3>> qp(ncorner);
  std::vector<double> qw(ncorner);
  const double p1[2] = {0.5 * (-1.0 / std::sqrt(3.0) + 1.0), 0.5 * (1.0 / std::sqrt(3.0) + 1.0)};
  for (int c = 0; c < ncorner; ++c) {
    double w = 1.0;
    for (int d = 0; d < dim; ++d) {
      int bit = (c >> (dim - 1 - d)) & 1;
      corners[c][d] = bit;
      qp[c][d] = p1[bit];
      w *= 0.5 * 1.0;
    }
    qw[c] = w;
  }
  long Ncell = lat.Ncell();
  for (long cell = 0; cell < Ncell; ++cell) {
    int cc[3];
    lat.c2e(cell, cc);
    bool overlap = false;
    double xmin[3] = {2, 2, 2}, xmax[3] = {-1, -1, -1};
    for (int c = 0; c < ncorner; ++c) {
      double s = 0;
      for (int d = 0; d < dim; ++d) {
        double xc = h[d] * double(cc[d] + corners[c][d]);
        xmin[d] = std::min(xmin[d], xc);
        xmax[d] = std::max(xmax[d], xc);
        s += (xc - x0[d]) * (xc - x0[d]);
      }
      overlap = overlap || (std::sqrt(s) < radius);
    }
    bool centre_in_cell = true;
    for (int d = 0; d < dim; ++d) centre_in_cell = centre_in_cell && (xmin[d] <= x0[d]) && (x0[d] <= xmax[d]);
    overlap = overlap || centre_in_cell;
    if (!overlap) continue;
    for (int a = 0; a < ncorner; ++a) {
      int vi[3];
      for (int d = 0; d < dim; ++d) vi[d] = cc[d] + corners[a][d];
      long ell = lat.e2v(vi);
      if (ell < 0) continue;
      double local = 0.0;
      for (int q = 0; q < ncorner; ++q) {
        double s = 0;
        for (int d = 0; d < dim; ++d) {
          double x = h[d] * (qp[q][d] + cc[d]);
          s += (x - x0[d]) * (x - x0[d]);
        }
        double xi = std::sqrt(s) / radius;
        if (xi < 1.0) {
          double phihat = 1.0;
          for (int d = 0; d < dim; ++d) phihat *= (corners[a][d] == 0) ? (1.0 - qp[q][d]) : qp[q][d];
          local += phihat * qw[q] * cv * normalisation;
        }
      }
      r[ell] += local;
    }
  }
  return r;
}

// MeasuredOperator ctor (measured_operator.cc:9-49)
static void add_measurements(Op& op, int nmeas, const double* centres, double radius, const double* variance,
                             double variance_scaling, int measure_global, double variance_global) {
  long N = op.N();
  op.m = nmeas + (measure_global ? 1 : 0);
  op.Sigma.assign(op.m, 0.0);
  std::vector<CSR::Trip> t;
  for (int k = 0; k < nmeas; ++k) {
    op.Sigma[k] = variance_scaling * variance[k];
    std::vector<double> r = measurement_vector(op.lat, centres + k * op.lat.dim, radius);
    for (long i = 0; i < N; ++i)
      if (r[i] != 0.0) t.push_back({i, k, r[i]});
  }
  if (measure_global) {
    double cv = op.lat.cell_volume();
    for (long i = 0; i < N; ++i) t.push_back({i, nmeas, cv});
    op.Sigma[nmeas] = variance_global;
  }
  op.B = CSR::from_triplets(N, op.m, t);
}

// ---------------------------------------------------------------------------
// Intergrid: src/intergrid/intergrid_operator{.hh,.cc,_linear.cc}
// ---------------------------------------------------------------------------
struct Intergrid {
  Lattice fine, coarse;
  int ssize = 0;
  std::vector<double> matrix;
  std::vector<long> colidx;
  explicit Intergrid(const Lattice& lat) : fine(lat), coarse(lat.coarse()) {
    int dim = lat.dim;
    ssize = 1;
    for (int d = 0; d < dim; ++d) ssize *= 3;
    const double st1[3] = {0.5, 1.0, 0.5};
    const int sh1[3] = {-1, 0, 1};
    matrix.assign(ssize, 1.0);
    std::vector<std::array<int, 3>> shift(ssize);
    for (int j = 0; j < ssize; ++j) {  // intergrid_operator_linear.cc:16-28
      int mu = j;
      for (int d = 0; d < dim; ++d) {
        matrix[j] *= st1[mu % 3];
        shift[j][d] = sh1[mu % 3];
        mu /= 3;
      }
    }
    long Nc = coarse.Nvertex();
    colidx.resize(Nc * ssize);
    for (long I = 0; I < Nc; ++I) {  // intergrid_operator.cc:8-20
      int ic[3];
      coarse.v2e(I, ic);
      for (int j = 0; j < ssize; ++j) {
        int f[3];
        for (int d = 0; d < dim; ++d) f[d] = 2 * ic[d] + shift[j][d];
        long e = fine.e2v(f);
        if (e < 0) throw std::runtime_error("intergrid shift leaves interior");
        colidx[I * ssize + j] = e;
      }
    }
  }
  void restrict_(const double* x, double* xc) const {  // intergrid_operator.hh:74-88
    long Nc = coarse.Nvertex();
    for (long I = 0; I < Nc; ++I) {
      double s = 0;
      for (int k = 0; k < ssize; ++k) s += matrix[k] * x[colidx[I * ssize + k]];
      xc[I] = s;
    }
  }
  void prolongate_add(double alpha, const double* xc, double* x) const {  // :106-120
    long Nc = coarse.Nvertex();
    for (long I = 0; I < Nc; ++I) {
      double v = xc[I];
      for (int k = 0; k < ssize; ++k) x[colidx[I * ssize + k]] += alpha * matrix[k] * v;
    }
  }
  CSR to_sparse() const {  // :123-144
    std::vector<CSR::Trip> t;
    long Nc = coarse.Nvertex();
    for (long I = 0; I < Nc; ++I)
      for (int k = 0; k < ssize; ++k) t.push_back({I, colidx[I * ssize + k], matrix[k]});
    return CSR::from_triplets(Nc, fine.Nvertex(), t);
  }
};

// LinearOperator::coarsen (linear_operator.cc:10-23)
static Op coarsen(const Op& op, const Intergrid& ig) {
  CSR R = ig.to_sparse();
  CSR P = R.transpose();
  Op c;
  c.lat = ig.coarse;
  c.A = R.mul(op.A).mul(P);
  c.m = op.m;
  c.Sigma = op.Sigma;
  if (op.m > 0) c.B = R.mul(op.B);
  else {
    c.B.nrow = c.A.nrow;
    c.B.ncol = 0;
    c.B.rp.assign(c.A.nrow + 1, 0);
  }
  return c;
}

// ---------------------------------------------------------------------------
// Orderings. Lexicographic = the reference (sor_smoother.cc:66-68).
// Multicolour = the HIP kernels' order: colour-major, ascending index within
// a colour. Star stencils (5/7-point): colour = (sum of vertex indices) & 1.
// Box stencils (9/27-point): colour = sum_d (idx_d & 1) << d.
// ---------------------------------------------------------------------------
enum { ORDER_LEX = 0, ORDER_MULTICOLOUR = 1 };

static bool is_star_stencil(const Op& op) {
  const Lattice& L = op.lat;
  for (long r = 0; r < op.A.nrow; ++r) {
    int a[3], b[3];
    L.v2e(r, a);
    for (long k = op.A.rp[r]; k < op.A.rp[r + 1]; ++k) {
      L.v2e(op.A.ci[k], b);
      int nz = 0;
      for (int d = 0; d < L.dim; ++d) {
        int o = b[d] - a[d];
        if (o < -1 || o > 1) throw std::runtime_error("stencil wider than nearest neighbours: no multicolour ordering");
        nz += (o != 0);
      }
      if (nz > 1) return false;
    }
  }
  return true;
}

static std::vector<long> make_order(const Op& op, int ordering, bool forward) {
  long N = op.N();
  std::vector<long> ord(N);
  if (ordering == ORDER_LEX) {
    for (long i = 0; i < N; ++i) ord[i] = i;
  } else {
    bool star = is_star_stencil(op);
    int ncol = star ? 2 : (1 << op.lat.dim);
    std::vector<int> colour(N);
    for (long i = 0; i < N; ++i) {
      int idx[3];
      op.lat.v2e(i, idx);
      int c = 0;
      if (star) {
        for (int d = 0; d < op.lat.dim; ++d) c += idx[d];
        c &= 1;
      } else {
        for (int d = 0; d < op.lat.dim; ++d) c |= (idx[d] & 1) << d;
      }
      colour[i] = c;
    }
    long p = 0;
    for (int c = 0; c < ncol; ++c)
      for (long i = 0; i < N; ++i)
        if (colour[i] == c) ord[p++] = i;
  }
  if (!forward) std::reverse(ord.begin(), ord.end());
  return ord;
}

// ---------------------------------------------------------------------------
// SOR smoother with SMW low-rank correction: src/smoother/sor_smoother.cc
// ---------------------------------------------------------------------------
struct SORSmoother {
  const Op* op;
  double omega;
  int nsmooth;
  std::vector<long> order;      // visiting order of the sweep
  std::vector<long> position;   // inverse permutation
  std::vector<double> Bbar;     // dense N x m, row-major
  std::vector<double> diag;

  SORSmoother(const Op* op_, double omega_, int nsmooth_, bool forward, int ordering)
      : op(op_), omega(omega_), nsmooth(nsmooth_) {
    order = make_order(*op, ordering, forward);
    long N = op->N();
    position.resize(N);
    for (long p = 0; p < N; ++p) position[order[p]] = p;
    diag.resize(N);
    for (long i = 0; i < N; ++i) diag[i] = op->A.diag(i);
    int m = op->m;
    if (m > 0) {  // sor_smoother.cc:17-37
      // W = (L_order + D/omega)^{-1} B, one column at a time
      std::vector<double> W(N * m, 0.0), col(N), w(N);
      for (int k = 0; k < m; ++k) {
        std::fill(col.begin(), col.end(), 0.0);
        for (long i = 0; i < N; ++i)
          for (long a = op->B.rp[i]; a < op->B.rp[i + 1]; ++a)
            if (op->B.ci[a] == k) col[i] += op->B.v[a];
        for (long p = 0; p < N; ++p) {
          long i = order[p];
          double s = col[i];
          for (long a = op->A.rp[i]; a < op->A.rp[i + 1]; ++a) {
            long j = op->A.ci[a];
            if (position[j] < p) s -= op->A.v[a] * w[j];
          }
          w[i] = s / (diag[i] / omega);
        }
        for (long i = 0; i < N; ++i) W[i * m + k] = w[i];
      }
      // S = Sigma + B^T W ; Bbar = W S^{-1}
      std::vector<double> S(m * m, 0.0);
      for (long i = 0; i < N; ++i)
        for (long a = op->B.rp[i]; a < op->B.rp[i + 1]; ++a) {
          int kk = (int)op->B.ci[a];
          for (int l = 0; l < m; ++l) S[kk * m + l] += op->B.v[a] * W[i * m + l];
        }
      for (int k = 0; k < m; ++k) S[k * m + k] += op->Sigma[k];
      std::vector<double> Sinv = inverse(S, m);
      Bbar.assign(N * m, 0.0);
      for (long i = 0; i < N; ++i)
        for (int l = 0; l < m; ++l) {
          double s = 0;
          for (int k = 0; k < m; ++k) s += W[i * m + k] * Sinv[k * m + l];
          Bbar[i * m + l] = s;
        }
    }
  }
  // SORSmoother::apply_sparse (sor_smoother.cc:56-77). Note the reference
  // loops nsmooth times here AND in apply(), so apply() with nsmooth = n
  // performs n*n sparse sweeps; this is kept.
  void apply_sparse(const double* b, double* x) const {
    const CSR& A = op->A;
    long N = op->N();
    for (int s = 0; s < nsmooth; ++s)
      for (long p = 0; p < N; ++p) {
        long i = order[p];
        double res = 0.0;
        for (long a = A.rp[i]; a < A.rp[i + 1]; ++a) res += A.v[a] * x[A.ci[a]];
        x[i] += omega * (b[i] - res) / diag[i];
      }
  }
  void lowrank(double* x) const {  // sor_smoother.cc:46-51
    int m = op->m;
    long N = op->N();
    std::vector<double> z(m);
    op->B.matvec_T(x, z.data());
    for (long i = 0; i < N; ++i) {
      double s = 0;
      for (int k = 0; k < m; ++k) s += Bbar[i * m + k] * z[k];
      x[i] -= s;
    }
  }
  void apply(const double* b, double* x) const {  // sor_smoother.cc:41-53
    for (int k = 0; k < nsmooth; ++k) {
      apply_sparse(b, x);
      if (op->m > 0) lowrank(x);
    }
  }
};

// SSORSmoother::apply (ssor_smoother.cc:9-15)
struct SSORSmoother {
  int nsmooth;
  std::shared_ptr<SORSmoother> fw, bw;
  void apply(const double* b, double* x) const {
    for (int k = 0; k < nsmooth; ++k) {
      fw->apply(b, x);
      bw->apply(b, x);
    }
  }
};

// ---------------------------------------------------------------------------
// Samplers: src/sampler/{sor,ssor}_sampler.cc, cholesky_sampler.{hh,cc}
// A "noise source" abstracts the RNG so tests can inject exact normals.
// ---------------------------------------------------------------------------
struct Noise {
  CLCG* rng = nullptr;
  const double* buf = nullptr;  // injected normals, consumed in order
  long pos = 0;
  double next() {
    if (rng) return rng->draw_normal();
    if (!buf) { ++pos; return 0.0; }  // counting mode
    return buf[pos++];
  }
};

struct SORSampler {
  const Op* op;
  double omega;
  int nsmooth;
  std::shared_ptr<SORSmoother> smoother;  // nsmooth = 1 (sor_sampler.cc:30)
  std::vector<double> sqrt_diag;
  std::vector<double> sigma_inv_sqrt;
  SORSampler(const Op* op_, double omega_, int nsmooth_, std::shared_ptr<SORSmoother> sm)
      : op(op_), omega(omega_), nsmooth(nsmooth_), smoother(sm) {
    long N = op->N();
    sqrt_diag.resize(N);
    for (long i = 0; i < N; ++i) sqrt_diag[i] = std::sqrt(op->A.diag(i) * (2. - omega) / omega);
    for (int k = 0; k < op->m; ++k) sigma_inv_sqrt.push_back(std::sqrt(1.0 / op->Sigma[k]));
  }
  // SORSampler::apply (sor_sampler.cc:37-59)
  void apply(const double* f, double* x, Noise& rng, std::vector<double>& c) const {
    long N = op->N();
    int m = op->m;
    c.resize(N);
    for (int k = 0; k < nsmooth; ++k) {
      for (long i = 0; i < N; ++i) {
        double t = sqrt_diag[i];
        c[i] = t * rng.next() + f[i];
      }
      if (m > 0) {
        std::vector<double> xi(m);
        for (int l = 0; l < m; ++l) xi[l] = rng.next() * sigma_inv_sqrt[l];
        for (long r = 0; r < N; ++r)
          for (long a = op->B.rp[r]; a < op->B.rp[r + 1]; ++a) c[r] += op->B.v[a] * xi[op->B.ci[a]];
      }
      smoother->apply(c.data(), x);
    }
  }
};

struct SSORSampler {
  int nsmooth;
  std::shared_ptr<SORSampler> fw, bw;
  void apply(const double* f, double* x, Noise& rng, std::vector<double>& c) const {
    for (int k = 0; k < nsmooth; ++k) {
      fw->apply(f, x, rng, c);
      bw->apply(f, x, rng, c);
    }
  }
};

// Dense Cholesky of Q = A + B Sigma^{-1} B^T for the coarsest level
// (cholesky_sampler.cc:9-23 with a dense factor in place of the AMD-permuted
// sparse one: identical distribution, different xi -> x map).
struct CholeskySampler {
  long n = 0;
  std::vector<double> L;
  explicit CholeskySampler(const Op& op) {
    n = op.N();
    L = op.dense_precision();
    if (!cholesky(L, n)) throw std::runtime_error("coarse precision matrix not SPD");
  }
  // CholeskySampler::apply (cholesky_sampler.hh:52-66)
  void apply(const double* f, double* x, Noise& rng) const {
    std::vector<double> xi(n), g(n);
    for (long i = 0; i < n; ++i) xi[i] = rng.next();
    solve_L(L, n, f, g.data());
    for (long i = 0; i < n; ++i) g[i] += xi[i];
    solve_LT(L, n, g.data(), x);
  }
};

// CholeskySolver (cholesky_solver.cc:8-40): A^{-1} by Cholesky, low rank by SMW
struct CholeskySolver {
  const Op* op;
  long n = 0;
  std::vector<double> L, Bbar;
  explicit CholeskySolver(const Op* op_) : op(op_) {
    n = op->N();
    L.assign(n * n, 0.0);
    for (long r = 0; r < n; ++r)
      for (long k = op->A.rp[r]; k < op->A.rp[r + 1]; ++k) L[r * n + op->A.ci[k]] += op->A.v[k];
    if (!cholesky(L, n)) throw std::runtime_error("coarse operator not SPD");
    int m = op->m;
    if (m > 0) {
      std::vector<double> AinvB(n * m), col(n), y(n), tmp(n);
      for (int k = 0; k < m; ++k) {
        std::fill(col.begin(), col.end(), 0.0);
        for (long i = 0; i < n; ++i)
          for (long a = op->B.rp[i]; a < op->B.rp[i + 1]; ++a)
            if (op->B.ci[a] == k) col[i] += op->B.v[a];
        solve_L(L, n, col.data(), tmp.data());
        solve_LT(L, n, tmp.data(), y.data());
        for (long i = 0; i < n; ++i) AinvB[i * m + k] = y[i];
      }
      std::vector<double> S(m * m, 0.0);
      for (long i = 0; i < n; ++i)
        for (long a = op->B.rp[i]; a < op->B.rp[i + 1]; ++a)
          for (int l = 0; l < m; ++l) S[op->B.ci[a] * m + l] += op->B.v[a] * AinvB[i * m + l];
      for (int k = 0; k < m; ++k) S[k * m + k] += op->Sigma[k];
      std::vector<double> Sinv = inverse(S, m);
      Bbar.assign(n * m, 0.0);
      for (long i = 0; i < n; ++i)
        for (int l = 0; l < m; ++l) {
          double s = 0;
          for (int k = 0; k < m; ++k) s += AinvB[i * m + k] * Sinv[k * m + l];
          Bbar[i * m + l] = s;
        }
    }
  }
  void apply(const double* b, double* x) const {
    std::vector<double> tmp(n), y(n);
    solve_L(L, n, b, tmp.data());
    solve_LT(L, n, tmp.data(), y.data());
    int m = op->m;
    if (m > 0) {
      std::vector<double> z(m);
      op->B.matvec_T(y.data(), z.data());
      for (long i = 0; i < n; ++i) {
        double s = 0;
        for (int k = 0; k < m; ++k) s += Bbar[i * m + k] * z[k];
        x[i] = y[i] - s;
      }
    } else {
      for (long i = 0; i < n; ++i) x[i] = y[i];
    }
  }
};

// ---------------------------------------------------------------------------
// Multigrid hierarchy shared by MultigridMCSampler
// (src/sampler/multigridmc_sampler.cc:8-145) and MultigridPreconditioner
// (src/preconditioner/multigrid_preconditioner.cc:8-109).
// ---------------------------------------------------------------------------
struct Params {
  int nlevel;
  int smoother;        // 0 = SOR, 1 = SSOR
  int npresmooth;
  int npostsmooth;
  int ncoarsesmooth;
  double omega;
  int cycle;
  double coarse_scaling;
  int coarse_solver;   // 0 = Cholesky, 1 = SSOR (sampler only)
  int ordering;        // ORDER_LEX or ORDER_MULTICOLOUR
};

struct Level {
  Op op;
  std::unique_ptr<Intergrid> ig;
  // smoothers (deterministic): pre/post, used by the MG preconditioner
  std::shared_ptr<SORSmoother> sor_fw_pre, sor_bw_post;          // "SOR" with n = npre/npost (quirk)
  std::shared_ptr<SORSmoother> sor_fw1, sor_bw1;                 // nsmooth = 1 (SSOR and samplers)
  std::shared_ptr<SORSampler> samp_fw_pre, samp_bw_post;         // "SOR" sampler n = npre/npost
  std::shared_ptr<SORSampler> samp_fw1, samp_bw1;                // nsmooth = 1
};

struct Hierarchy {
  Params p;
  std::vector<std::unique_ptr<Level>> levels;
  std::unique_ptr<CholeskySampler> coarse_chol;
  std::unique_ptr<CholeskySolver> coarse_solver;

  Hierarchy(const Op& fine, const Params& p_) : p(p_) {
    if (p.nlevel < 1) throw std::runtime_error("nlevel must be >= 1");
    if (p.smoother != 0 && p.smoother != 1) throw std::runtime_error("invalid smoother");
    Op cur = fine;
    for (int l = 0; l < p.nlevel; ++l) {
      auto L = std::make_unique<Level>();
      L->op = cur;
      const Op* op = &L->op;
      L->sor_fw1 = std::make_shared<SORSmoother>(op, p.omega, 1, true, p.ordering);
      L->sor_bw1 = std::make_shared<SORSmoother>(op, p.omega, 1, false, p.ordering);
      if (p.smoother == 0) {
        L->sor_fw_pre = (p.npresmooth == 1) ? L->sor_fw1
                                            : std::make_shared<SORSmoother>(op, p.omega, p.npresmooth, true, p.ordering);
        L->sor_bw_post = (p.npostsmooth == 1)
                             ? L->sor_bw1
                             : std::make_shared<SORSmoother>(op, p.omega, p.npostsmooth, false, p.ordering);
      }
      L->samp_fw1 = std::make_shared<SORSampler>(op, p.omega, 1, L->sor_fw1);
      L->samp_bw1 = std::make_shared<SORSampler>(op, p.omega, 1, L->sor_bw1);
      L->samp_fw_pre = std::make_shared<SORSampler>(op, p.omega, p.npresmooth, L->sor_fw1);
      L->samp_bw_post = std::make_shared<SORSampler>(op, p.omega, p.npostsmooth, L->sor_bw1);
      if (l < p.nlevel - 1) {
        L->ig = std::make_unique<Intergrid>(cur.lat);
        cur = coarsen(cur, *L->ig);
      }
      levels.push_back(std::move(L));
    }
    const Op& cop = levels.back()->op;
    coarse_chol = std::make_unique<CholeskySampler>(cop);
    coarse_solver = std::make_unique<CholeskySolver>(&cop);
  }
};

// per-chain scratch (the reference's mutable x_ell / f_ell / r_ell)
struct Work {
  std::vector<std::vector<double>> x, f, r;
  std::vector<double> c;
  explicit Work(const Hierarchy& H) {
    for (auto& L : H.levels) {
      long N = L->op.N();
      x.emplace_back(N, 0.0);
      f.emplace_back(N, 0.0);
      r.emplace_back(N, 0.0);
    }
  }
};

static void presample(const Hierarchy& H, int l, Work& w, Noise& rng) {
  const Level& L = *H.levels[l];
  if (H.p.smoother == 1) {
    SSORSampler s{H.p.npresmooth, L.samp_fw1, L.samp_bw1};
    s.apply(w.f[l].data(), w.x[l].data(), rng, w.c);
  } else {
    L.samp_fw_pre->apply(w.f[l].data(), w.x[l].data(), rng, w.c);
  }
}
static void postsample(const Hierarchy& H, int l, Work& w, Noise& rng) {
  const Level& L = *H.levels[l];
  if (H.p.smoother == 1) {
    SSORSampler s{H.p.npostsmooth, L.samp_fw1, L.samp_bw1};
    s.apply(w.f[l].data(), w.x[l].data(), rng, w.c);
  } else {
    L.samp_bw_post->apply(w.f[l].data(), w.x[l].data(), rng, w.c);
  }
}

// MultigridMCSampler::sample (multigridmc_sampler.cc:109-136)
static void mgmc_sample(const Hierarchy& H, int l, Work& w, Noise& rng) {
  if (l == H.p.nlevel - 1) {
    const Level& L = *H.levels[l];
    if (H.p.coarse_solver == 0) {
      H.coarse_chol->apply(w.f[l].data(), w.x[l].data(), rng);
    } else {
      SSORSampler s{H.p.ncoarsesmooth, L.samp_fw1, L.samp_bw1};
      s.apply(w.f[l].data(), w.x[l].data(), rng, w.c);
    }
    return;
  }
  const Level& L = *H.levels[l];
  int cyc = (l > 0) ? H.p.cycle : 1;
  for (int j = 0; j < cyc; ++j) {
    presample(H, l, w, rng);
    L.op.apply(w.x[l].data(), w.r[l].data());
    long N = L.op.N();
    for (long i = 0; i < N; ++i) w.r[l][i] = w.f[l][i] - w.r[l][i];
    L.ig->restrict_(w.r[l].data(), w.f[l + 1].data());
    std::fill(w.x[l + 1].begin(), w.x[l + 1].end(), 0.0);
    mgmc_sample(H, l + 1, w, rng);
    L.ig->prolongate_add(H.p.coarse_scaling, w.x[l + 1].data(), w.x[l].data());
    postsample(H, l, w, rng);
  }
}

static void presmooth(const Hierarchy& H, int l, Work& w) {
  const Level& L = *H.levels[l];
  if (H.p.smoother == 1) {
    SSORSmoother s{H.p.npresmooth, L.sor_fw1, L.sor_bw1};
    s.apply(w.f[l].data(), w.x[l].data());
  } else {
    L.sor_fw_pre->apply(w.f[l].data(), w.x[l].data());
  }
}
static void postsmooth(const Hierarchy& H, int l, Work& w) {
  const Level& L = *H.levels[l];
  if (H.p.smoother == 1) {
    SSORSmoother s{H.p.npostsmooth, L.sor_fw1, L.sor_bw1};
    s.apply(w.f[l].data(), w.x[l].data());
  } else {
    L.sor_bw_post->apply(w.f[l].data(), w.x[l].data());
  }
}

// MultigridPreconditioner::solve (multigrid_preconditioner.cc:74-101);
// w.f holds b_ell.
static void mg_solve(const Hierarchy& H, int l, Work& w) {
  std::fill(w.x[l].begin(), w.x[l].end(), 0.0);
  if (l == H.p.nlevel - 1) {
    H.coarse_solver->apply(w.f[l].data(), w.x[l].data());
    return;
  }
  const Level& L = *H.levels[l];
  int cyc = (l > 0) ? H.p.cycle : 1;
  for (int j = 0; j < cyc; ++j) {
    presmooth(H, l, w);
    L.op.apply(w.x[l].data(), w.r[l].data());
    long N = L.op.N();
    for (long i = 0; i < N; ++i) w.r[l][i] = w.f[l][i] - w.r[l][i];
    L.ig->restrict_(w.r[l].data(), w.f[l + 1].data());
    mg_solve(H, l + 1, w);
    L.ig->prolongate_add(H.p.coarse_scaling, w.x[l + 1].data(), w.x[l].data());
    postsmooth(H, l, w);
  }
}

// Statistics::tau_int restricted to a scalar series
// (src/auxilliary/statistics.cc:4-79): running averages as record_sample.
static double tau_int_reference(const double* q, long n, int window) {
  double avg = 0, avg2 = 0;
  std::vector<double> Sk;
  std::vector<double> hist;  // most recent first
  for (long s = 1; s <= n; ++s) {
    double Q = q[s - 1];
    if (s == 1) { avg = Q; avg2 = Q * Q; }
    else { avg += (Q - avg) / (1.0 * s); avg2 += (Q * Q - avg2) / (1.0 * s); }
    hist.insert(hist.begin(), Q);
    if ((int)hist.size() > window) hist.pop_back();
    for (unsigned k = 0; k < hist.size(); ++k) {
      long Nk = s - k;
      if (Nk == 1) Sk.push_back(hist[0] * hist[k]);
      else Sk[k] += (hist[0] * hist[k] - Sk[k]) / (1.0 * Nk);
    }
  }
  double var = Sk[0] - avg * avg;
  double tau = 1.0;
  unsigned kmax = Sk.size();
  for (unsigned k = 1; k < kmax; ++k) tau += 2 * (1. - k / (1.0 * kmax)) * (Sk[k] - avg * avg) / var;
  return tau;
}

}  // namespace orc

// ===========================================================================
// C ABI (ctypes)
// ===========================================================================
using namespace orc;

static thread_local std::string g_err;

struct orc_op { Op op; };
struct orc_mg { std::unique_ptr<Hierarchy> H; std::unique_ptr<Work> w; CLCG rng; };

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// prior: ShiftedLaplaceFDOperator with constant kappa^2
orc_op* orc_op_fd(int dim, int nx, int ny, int nz, double kappa_sq) {
  try {
    Lattice L;
    L.dim = dim;
    L.n[0] = nx;
    L.n[1] = ny;
    L.n[2] = nz;
    for (int d = 0; d < dim; ++d)
      if (L.n[d] < 2) throw std::runtime_error("lattice extent must be >= 2");
    auto* o = new orc_op;
    o->op = make_fd_operator(L, kappa_sq);
    return o;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// prior with the periodic correlation-length model
orc_op* orc_op_fd_periodic(int dim, int nx, int ny, int nz, double Lambda_min, double Lambda_max) {
  try {
    Lattice L;
    L.dim = dim;
    L.n[0] = nx;
    L.n[1] = ny;
    L.n[2] = nz;
    KappaModel k;
    k.periodic = true;
    k.L1 = 0.5 * (Lambda_max + Lambda_min);
    k.L2 = 0.5 * (Lambda_max - Lambda_min);
    auto* o = new orc_op;
    o->op = make_fd_operator(L, k);
    return o;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// operator from user CSR triplets (e.g. the reference's TestOperator1d)
orc_op* orc_op_custom(int dim, int nx, int ny, int nz, long nnz, const long* rows, const long* cols,
                      const double* vals) {
  try {
    Lattice L;
    L.dim = dim;
    L.n[0] = nx;
    L.n[1] = ny;
    L.n[2] = nz;
    long N = L.Nvertex();
    std::vector<CSR::Trip> t;
    for (long k = 0; k < nnz; ++k) t.push_back({rows[k], cols[k], vals[k]});
    auto* o = new orc_op;
    o->op.lat = L;
    o->op.A = CSR::from_triplets(N, N, t);
    o->op.B.nrow = N;
    o->op.B.rp.assign(N + 1, 0);
    return o;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

int orc_op_set_lowrank(orc_op* o, int m, long nnz, const long* rows, const long* cols, const double* vals,
                       const double* sigma) {
  try {
    std::vector<CSR::Trip> t;
    for (long k = 0; k < nnz; ++k) t.push_back({rows[k], cols[k], vals[k]});
    o->op.m = m;
    o->op.B = CSR::from_triplets(o->op.N(), m, t);
    o->op.Sigma.assign(sigma, sigma + m);
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return 1; }
}

int orc_op_add_measurements(orc_op* o, int nmeas, const double* centres, double radius, const double* variance,
                            double variance_scaling, int measure_global, double variance_global) {
  try {
    add_measurements(o->op, nmeas, centres, radius, variance, variance_scaling, measure_global, variance_global);
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return 1; }
}

void orc_op_free(orc_op* o) { delete o; }
long orc_op_ndof(const orc_op* o) { return o->op.N(); }
int orc_op_m(const orc_op* o) { return o->op.m; }
long orc_op_nnz(const orc_op* o, int which) { return which == 0 ? (long)o->op.A.v.size() : (long)o->op.B.v.size(); }
// export A (which=0) or B (which=1) as CSR
void orc_op_csr(const orc_op* o, int which, long* rp, long* ci, double* v) {
  const CSR& M = which == 0 ? o->op.A : o->op.B;
  std::copy(M.rp.begin(), M.rp.end(), rp);
  std::copy(M.ci.begin(), M.ci.end(), ci);
  std::copy(M.v.begin(), M.v.end(), v);
}
void orc_op_sigma(const orc_op* o, double* s) { std::copy(o->op.Sigma.begin(), o->op.Sigma.end(), s); }
void orc_op_apply(const orc_op* o, const double* x, double* y) { o->op.apply(x, y); }
void orc_op_dense_precision(const orc_op* o, double* Q) {
  std::vector<double> q = o->op.dense_precision();
  std::copy(q.begin(), q.end(), Q);
}
orc_op* orc_op_coarsen(const orc_op* o) {
  try {
    Intergrid ig(o->op.lat);
    auto* c = new orc_op;
    c->op = coarsen(o->op, ig);
    return c;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}
int orc_restrict(const orc_op* o, const double* x, double* xc) {
  try { Intergrid ig(o->op.lat); ig.restrict_(x, xc); return 0; }
  catch (std::exception& e) { g_err = e.what(); return 1; }
}
int orc_prolongate_add(const orc_op* o, double alpha, const double* xc, double* x) {
  try { Intergrid ig(o->op.lat); ig.prolongate_add(alpha, xc, x); return 0; }
  catch (std::exception& e) { g_err = e.what(); return 1; }
}
int orc_measurement_vector(const orc_op* o, const double* x0, double radius, double* out) {
  std::vector<double> r = measurement_vector(o->op.lat, x0, radius);
  std::copy(r.begin(), r.end(), out);
  return 0;
}
// Q^{-1} b by dense Cholesky (small problems only)
int orc_op_dense_solve(const orc_op* o, const double* b, double* x) {
  long n = o->op.N();
  std::vector<double> L = o->op.dense_precision(), t(n);
  if (!cholesky(L, n)) { g_err = "not SPD"; return 1; }
  solve_L(L, n, b, t.data());
  solve_LT(L, n, t.data(), x);
  return 0;
}

orc_mg* orc_mg_create(const orc_op* o, int nlevel, int smoother, int npre, int npost, int ncoarse, double omega,
                      int cycle, double coarse_scaling, int coarse_solver, int ordering) {
  try {
    Params p{nlevel, smoother, npre, npost, ncoarse, omega, cycle, coarse_scaling, coarse_solver, ordering};
    auto* g = new orc_mg;
    g->H = std::make_unique<Hierarchy>(o->op, p);
    g->w = std::make_unique<Work>(*g->H);
    return g;
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}
void orc_mg_free(orc_mg* g) { delete g; }
int orc_mg_nlevel(const orc_mg* g) { return (int)g->H->levels.size(); }
long orc_mg_level_ndof(const orc_mg* g, int l) { return g->H->levels[l]->op.N(); }
// export the Galerkin operator of level l (which = 0: A, 1: B)
long orc_mg_level_nnz(const orc_mg* g, int l, int which) {
  const Op& op = g->H->levels[l]->op;
  return which == 0 ? (long)op.A.v.size() : (long)op.B.v.size();
}
void orc_mg_level_csr(const orc_mg* g, int l, int which, long* rp, long* ci, double* v) {
  const Op& op = g->H->levels[l]->op;
  const CSR& M = which == 0 ? op.A : op.B;
  std::copy(M.rp.begin(), M.rp.end(), rp);
  std::copy(M.ci.begin(), M.ci.end(), ci);
  std::copy(M.v.begin(), M.v.end(), v);
}
void orc_mg_level_apply(const orc_mg* g, int l, const double* x, double* y) { g->H->levels[l]->op.apply(x, y); }
// Bbar (N x m, row-major) of the level's nsmooth=1 SOR smoother
void orc_mg_level_bbar(const orc_mg* g, int l, int forward, double* out) {
  const Level& L = *g->H->levels[l];
  const auto& s = forward ? L.sor_fw1 : L.sor_bw1;
  std::copy(s->Bbar.begin(), s->Bbar.end(), out);
}

// one deterministic SOR sweep (+SMW) of level l: SORSmoother::apply, nsmooth=1
void orc_mg_sor_smoother(const orc_mg* g, int l, int forward, const double* b, double* x) {
  const Level& L = *g->H->levels[l];
  (forward ? L.sor_fw1 : L.sor_bw1)->apply(b, x);
}
// one SOR Gibbs sweep of level l with injected normals: xi[0..N) then xi_lr[0..m)
void orc_mg_sor_sampler_noise(const orc_mg* g, int l, int forward, const double* f, double* x, const double* xi,
                              const double* xi_lr) {
  const Level& L = *g->H->levels[l];
  long N = L.op.N();
  int m = L.op.m;
  std::vector<double> buf(N + m);
  std::copy(xi, xi + N, buf.begin());
  if (m > 0) std::copy(xi_lr, xi_lr + m, buf.begin() + N);
  Noise nz;
  nz.buf = buf.data();
  std::vector<double> c;
  (forward ? L.samp_fw1 : L.samp_bw1)->apply(f, x, nz, c);
}
// coarsest-level Cholesky sample with injected normals
void orc_mg_coarse_sample_noise(const orc_mg* g, const double* f, double* x, const double* xi) {
  Noise nz;
  nz.buf = xi;
  g->H->coarse_chol->apply(f, x, nz);
}
void orc_mg_coarse_solve(const orc_mg* g, const double* b, double* x) { g->H->coarse_solver->apply(b, x); }
int orc_mg_restrict(const orc_mg* g, int l, const double* x, double* xc) {
  if (!g->H->levels[l]->ig) return 1;
  g->H->levels[l]->ig->restrict_(x, xc);
  return 0;
}
int orc_mg_prolongate_add(const orc_mg* g, int l, double alpha, const double* xc, double* x) {
  if (!g->H->levels[l]->ig) return 1;
  g->H->levels[l]->ig->prolongate_add(alpha, xc, x);
  return 0;
}

// MultigridPreconditioner::apply (multigrid_preconditioner.cc:104-109)
void orc_mg_precondition(orc_mg* g, const double* b, double* x) {
  Work& w = *g->w;
  std::copy(b, b + w.f[0].size(), w.f[0].begin());
  mg_solve(*g->H, 0, w);
  std::copy(w.x[0].begin(), w.x[0].end(), x);
}

// LoopSolver::apply (loop_solver.cc:9-54). hist[k] = ||r_k||; returns the
// number of iterations (index of the converged residual), -1 if not converged.
int orc_loop_solve(orc_mg* g, const double* b, double* x, double rtol, double atol, int maxiter, double* hist) {
  const Op& op = g->H->levels[0]->op;
  long N = op.N();
  double r0 = 0;
  for (long i = 0; i < N; ++i) r0 += b[i] * b[i];
  r0 = std::sqrt(r0);
  std::fill(x, x + N, 0.0);
  std::vector<double> r(N), Pr(N);
  for (int k = 0; k < maxiter; ++k) {
    op.apply(x, r.data());
    double rn = 0;
    for (long i = 0; i < N; ++i) { r[i] -= b[i]; rn += r[i] * r[i]; }
    rn = std::sqrt(rn);
    if (hist) hist[k] = rn;
    if ((rn / r0 < rtol) && (rn < atol)) return k;
    orc_mg_precondition(g, r.data(), Pr.data());
    for (long i = 0; i < N; ++i) x[i] -= Pr[i];
  }
  return -1;
}

// MultigridMCSampler::apply with the CLCG generator (one chain)
void orc_mgmc_reset_rng(orc_mg* g) { g->rng = CLCG(); }
void orc_mgmc_apply(orc_mg* g, const double* f, double* x) {
  Work& w = *g->w;
  std::copy(f, f + w.f[0].size(), w.f[0].begin());
  std::copy(x, x + w.x[0].size(), w.x[0].begin());
  Noise nz;
  nz.rng = &g->rng;
  mgmc_sample(*g->H, 0, w, nz);
  std::copy(w.x[0].begin(), w.x[0].end(), x);
}

// MultigridMCSampler::apply with injected normals, consumed in the order the
// reference draws them; xi == NULL only counts the draws (returned).
long orc_mgmc_apply_noise(orc_mg* g, const double* f, double* x, const double* xi) {
  Work& w = *g->w;
  std::copy(f, f + w.f[0].size(), w.f[0].begin());
  std::copy(x, x + w.x[0].size(), w.x[0].begin());
  Noise nz;
  nz.buf = xi;
  mgmc_sample(*g->H, 0, w, nz);
  std::copy(w.x[0].begin(), w.x[0].end(), x);
  return nz.pos;
}

// Independent chains, one per OpenMP thread, each with its own leapfrogged
// CLCG stream and scratch (the pattern of src/measurements/convergence.cc:29-63).
// x: nchains x N (in/out). obs: nsamples x nchains x nprobe, the values of x
// at the probe vertices after each sample (may be NULL). Returns wall seconds
// of the sampling loop.
double orc_mgmc_chains(orc_mg* g, const double* f, double* x, int nchains, int nsamples, int nthreads,
                       int nprobe, const long* probes, double* obs) {
  const Hierarchy& H = *g->H;
  long N = H.levels[0]->op.N();
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int c = 0; c < nchains; ++c) {
    Work w(H);
    CLCG rng(nchains, c);
    Noise nz;
    nz.rng = &rng;
    std::copy(f, f + N, w.f[0].begin());
    std::copy(x + (long)c * N, x + (long)(c + 1) * N, w.x[0].begin());
    for (int s = 0; s < nsamples; ++s) {
      std::copy(f, f + N, w.f[0].begin());
      mgmc_sample(H, 0, w, nz);
      if (obs)
        for (int p = 0; p < nprobe; ++p) obs[((long)s * nchains + c) * nprobe + p] = w.x[0][probes[p]];
    }
    std::copy(w.x[0].begin(), w.x[0].end(), x + (long)c * N);
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// CLCG stream access for the RNG known-answer tests (test_random.hh)
struct orc_rng { CLCG g; };
orc_rng* orc_rng_create(int nstreams, int id) { auto* r = new orc_rng{CLCG(nstreams, id)}; return r; }
void orc_rng_free(orc_rng* r) { delete r; }
void orc_rng_ints(orc_rng* r, long n, int32_t* out) { for (long i = 0; i < n; ++i) out[i] = r->g.draw_int(); }
void orc_rng_uniform(orc_rng* r, long n, double* out) { for (long i = 0; i < n; ++i) out[i] = r->g.draw_uniform(); }
void orc_rng_normal(orc_rng* r, long n, double* out) { for (long i = 0; i < n; ++i) out[i] = r->g.draw_normal(); }

double orc_tau_int_reference(const double* q, long n, int window) { return tau_int_reference(q, n, window); }

// The right-hand side of driver_mg (src/driver_mg.cc:163-171): n draws of
// std::normal_distribution<double>(0, 1) from std::mt19937_64(1482817), with
// the standard library the drivers are built against.
void orc_driver_mg_rhs(long n, double* out) {
  std::mt19937_64 rng(1482817);
  std::normal_distribution<double> normal_dist(0.0, 1.0);
  for (long i = 0; i < n; ++i) out[i] = normal_dist(rng);
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
