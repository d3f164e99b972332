// Implementation of the adapter classes (mgmc_adapters.hh) over the C ABI.
#include "mgmc_adapters.hh"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace mgmc_amd {

static void check(int status, const char* what) {
  if (status != MGMC_OK) throw Error(status, std::string(what) + ": " + mgmc_last_error());
}

int64_t Lattice::Nvertex() const {
  int64_t N = 1;
  for (int d = 0; d < dim; ++d) N *= n[d] - 1;
  return N;
}

std::shared_ptr<LinearOperator> LinearOperator::shifted_laplace_fd(const Lattice& lattice, double kappa_sq) {
  auto op = std::make_shared<LinearOperator>();
  op->lattice = lattice;
  op->kappa_sq = kappa_sq;
  return op;
}

std::shared_ptr<LinearOperator> LinearOperator::from_csr(const Lattice& lattice, std::vector<int64_t> rowptr,
                                                         std::vector<int64_t> col, std::vector<double> val) {
  if ((int64_t)rowptr.size() != lattice.Nvertex() + 1) throw Error(MGMC_ERR_ARG, "A: rowptr has the wrong length");
  auto op = std::make_shared<LinearOperator>();
  op->lattice = lattice;
  op->A_rowptr = std::move(rowptr);
  op->A_col = std::move(col);
  op->A_val = std::move(val);
  return op;
}

std::shared_ptr<LinearOperator> LinearOperator::measured(const std::shared_ptr<LinearOperator>& prior,
                                                         const std::vector<double>& centres, double radius,
                                                         const std::vector<double>& variance,
                                                         double variance_scaling, int profile, double width) {
  const Lattice& lat = prior->lattice;
  const int m = (int)variance.size();
  if ((int)centres.size() != m * lat.dim) throw Error(MGMC_ERR_ARG, "measured: centres must hold m x dim values");
  auto op = std::make_shared<LinearOperator>(*prior);
  const int64_t N = lat.Nvertex();
  // columns of B (MeasuredOperator::measurement_vector), then CSR rows
  std::vector<std::vector<std::pair<int, double>>> rows((size_t)N);
  std::vector<int64_t> idx;
  std::vector<double> val;
  for (int k = 0; k < m; ++k) {
    int64_t nnz = 0;
    int64_t cap = 1024;
    for (;;) {
      idx.resize((size_t)cap);
      val.resize((size_t)cap);
      const int st = mgmc_measurement_vector_weighted(lat.dim, lat.n, centres.data() + (size_t)k * lat.dim, radius,
                                                      profile, width, cap, idx.data(), val.data(), &nnz);
      if (st == MGMC_OK) break;
      if (st == MGMC_ERR_ARG && nnz > cap) {
        cap = nnz;
        continue;
      }
      check(st, "mgmc_measurement_vector");
    }
    for (int64_t e = 0; e < nnz; ++e) rows[(size_t)idx[e]].push_back({k, val[e]});
  }
  op->m = m;
  op->B_rowptr.assign(1, 0);
  op->B_col.clear();
  op->B_val.clear();
  for (auto& r : rows) {
    for (auto& e : r) {
      op->B_col.push_back(e.first);
      op->B_val.push_back(e.second);
    }
    op->B_rowptr.push_back((int64_t)op->B_col.size());
  }
  op->sigma.resize(m);
  for (int k = 0; k < m; ++k) op->sigma[k] = variance_scaling * variance[k];
  return op;
}

std::shared_ptr<LinearOperator> LinearOperator::with_measurements(const std::shared_ptr<LinearOperator>& prior, int m,
                                                                  std::vector<int64_t> B_rowptr,
                                                                  std::vector<int64_t> B_col, std::vector<double> B_val,
                                                                  std::vector<double> sigma) {
  if ((int64_t)B_rowptr.size() != prior->lattice.Nvertex() + 1 || (int)sigma.size() != m)
    throw Error(MGMC_ERR_ARG, "with_measurements: B must have N + 1 row pointers and Sigma m entries");
  auto op = std::make_shared<LinearOperator>(*prior);
  op->m = m;
  op->B_rowptr = std::move(B_rowptr);
  op->B_col = std::move(B_col);
  op->B_val = std::move(B_val);
  op->sigma = std::move(sigma);
  return op;
}

std::vector<double> LinearOperator::BT(ConstVec x) const {
  std::vector<double> z((size_t)m, 0.0);
  for (size_t r = 0; r + 1 < B_rowptr.size(); ++r)
    for (int64_t e = B_rowptr[r]; e < B_rowptr[r + 1]; ++e) z[(size_t)B_col[e]] += B_val[e] * x.p[r];
  return z;
}

std::vector<double> LinearOperator::B_Sigma_inv(ConstVec y) const {
  std::vector<double> f((size_t)lattice.Nvertex(), 0.0);
  for (size_t r = 0; r + 1 < B_rowptr.size(); ++r)
    for (int64_t e = B_rowptr[r]; e < B_rowptr[r + 1]; ++e)
      f[r] += B_val[e] * (y.p[(size_t)B_col[e]] / sigma[(size_t)B_col[e]]);
  return f;
}

mgmc_problem LinearOperator::problem() const {
  mgmc_problem p{};
  p.dim = lattice.dim;
  for (int d = 0; d < 3; ++d) p.n[d] = lattice.n[d];
  p.kappa_sq = kappa_sq;
  if (!A_rowptr.empty()) {
    p.A_rowptr = A_rowptr.data();
    p.A_col = A_col.data();
    p.A_val = A_val.data();
  }
  p.m = m;
  if (m > 0) {
    p.B_rowptr = B_rowptr.data();
    p.B_col = B_col.data();
    p.B_val = B_val.data();
    p.sigma = sigma.data();
  }
  return p;
}

Handle::Handle(const LinearOperator& op, const MultigridParameters& p, int coarse_solver, const DeviceOptions& o,
               uint64_t seed) {
  if (p.smoother != "SOR" && p.smoother != "SSOR") throw Error(MGMC_ERR_ARG, "invalid smoother '" + p.smoother + "'");
  mgmc_problem prob = op.problem();
  mgmc_params par{};
  par.nlevel = p.nlevel;
  par.smoother = p.smoother == "SSOR" ? MGMC_SSOR : MGMC_SOR;
  par.npresmooth = p.npresmooth;
  par.npostsmooth = p.npostsmooth;
  par.ncoarsesmooth = p.ncoarsesmooth;
  par.omega = p.omega;
  par.cycle = p.cycle;
  par.coarse_scaling = p.coarse_scaling;
  par.coarse_solver = coarse_solver;
  par.precision = o.precision;
  par.nchains = 1;
  par.seed = seed;
  par.stream = o.stream;
  check(mgmc_create(&prob, &par, &h), "mgmc_create");
  if (o.ordering != MGMC_ORDER_MULTICOLOUR) {
    const int st = mgmc_set_ordering(h, o.ordering);
    if (st != MGMC_OK) {
      mgmc_destroy(h);
      h = nullptr;
      check(st, "mgmc_set_ordering");
    }
  }
}

Handle::~Handle() {
  if (h) mgmc_destroy(h);
}

static uint64_t draw_key(RandomGenerator& rng) {
  const uint64_t hi = (uint32_t)rng.draw_int(), lo = (uint32_t)rng.draw_int();
  return (hi << 32) | lo;
}

static int coarse_of(const MultigridParameters& p) {
  if (p.coarse_solver == "Cholesky") return MGMC_COARSE_CHOLESKY;
  if (p.coarse_solver == "SSOR") return MGMC_COARSE_SSOR;
  throw Error(MGMC_ERR_ARG, "invalid coarse solver '" + p.coarse_solver + "'");
}

HipMultigridMCSampler::HipMultigridMCSampler(std::shared_ptr<LinearOperator> linear_operator_,
                                             std::shared_ptr<RandomGenerator> rng_, const MultigridParameters& p,
                                             const DeviceOptions& o)
    : Sampler(std::move(linear_operator_), std::move(rng_)), params(p), options(o) {
  if (!rng) throw Error(MGMC_ERR_ARG, "HipMultigridMCSampler needs a random generator");
  key = draw_key(*rng);
  h = std::make_unique<Handle>(*linear_operator, params, coarse_of(params), options, key);
}

std::shared_ptr<Sampler> HipMultigridMCSampler::deep_copy(std::shared_ptr<RandomGenerator> rng_) {
  return std::make_shared<HipMultigridMCSampler>(linear_operator, std::move(rng_), params, options);
}

void HipMultigridMCSampler::apply(ConstVec f, Vec x) const {
  const size_t N = (size_t)linear_operator->get_lattice().Nvertex();
  if (x.n != N) throw Error(MGMC_ERR_ARG, "Sampler::apply: x must hold one value per vertex");
  if (rhs_fixed) {
    check(mgmc_set_state(h->get(), x.p), "mgmc_set_state");
    check(mgmc_run(h->get(), 1, nullptr, 0), "mgmc_run");
    check(mgmc_get_state(h->get(), x.p), "mgmc_get_state");
    return;
  }
  if (f.n != N) throw Error(MGMC_ERR_ARG, "Sampler::apply: f must hold one value per vertex");
  check(mgmc_sample(h->get(), f.p, 0, x.p), "mgmc_sample");
}

void HipMultigridMCSampler::fix_rhs(ConstVec f) {
  if (f.n != (size_t)linear_operator->get_lattice().Nvertex())
    throw Error(MGMC_ERR_ARG, "fix_rhs: f must hold one value per vertex");
  check(mgmc_set_rhs(h->get(), f.p, 0), "mgmc_set_rhs");
  rhs_fixed = true;
}

HipMultigridPreconditioner::HipMultigridPreconditioner(std::shared_ptr<LinearOperator> linear_operator_,
                                                       const MultigridParameters& p, const DeviceOptions& o)
    : Preconditioner(std::move(linear_operator_)) {
  if (p.coarse_solver != "Cholesky")
    std::printf("WARNING: ignoring coarse solver setting '%s', using Cholesky.\n", p.coarse_solver.c_str());
  h = std::make_unique<Handle>(*linear_operator, p, MGMC_COARSE_CHOLESKY, o, 0);
}

void HipMultigridPreconditioner::apply(ConstVec b, Vec x) {
  const size_t N = (size_t)linear_operator->get_lattice().Nvertex();
  if (b.n != N || x.n != N) throw Error(MGMC_ERR_ARG, "Preconditioner::apply: b and x must hold N values");
  check(mg_precondition(h->get(), b.p, x.p), "mg_precondition");
}

HipLoopSolver::HipLoopSolver(std::shared_ptr<LinearOperator> linear_operator_,
                             std::shared_ptr<HipMultigridPreconditioner> preconditioner_,
                             const IterativeSolverParameters& params_)
    : linear_operator(std::move(linear_operator_)), preconditioner(std::move(preconditioner_)), params(params_) {}

void HipLoopSolver::apply(ConstVec b, Vec x) {
  const size_t N = (size_t)linear_operator->get_lattice().Nvertex();
  if (b.n != N || x.n != N) throw Error(MGMC_ERR_ARG, "LoopSolver::apply: b and x must hold N values");
  hist.assign((size_t)std::max(params.maxiter, 1), 0.0);
  int it = -1;
  const int st = mg_solve(preconditioner->handle(), b.p, x.p, params.rtol, params.atol, params.maxiter, hist.data(), &it);
  if (st != MGMC_OK && st != MGMC_ERR_NOCONV) check(st, "mg_solve");
  niter_ = it;
  hist.resize(it >= 0 ? (size_t)it + 1 : (size_t)params.maxiter);
  double r0 = 0.0;
  for (size_t i = 0; i < N; ++i) r0 += b.p[i] * b.p[i];
  r0 = std::sqrt(r0);
  if (params.verbose >= 2) {
    std::printf("Initial residual ||r_0|| =  %12.4f\n", r0);
    std::printf("%5s   %8s   %12s   %6s\n", "iter", "||r||", "||r||/||r_0||", "rho");
    double rold = r0;
    for (size_t k = 0; k < hist.size(); ++k) {
      std::printf("%5d   %8.3e   %12.3e   %6.3f\n", (int)k, hist[k], hist[k] / r0, hist[k] / rold);
      rold = hist[k];
    }
  }
  if (params.verbose >= 1) {
    if (niter_ >= 0)
      std::printf("Solver converged after %5d iterations\n||r|| = %8.3e, ||r||/||r_0|| = %8.3e\n", niter_,
                  hist[(size_t)niter_], hist[(size_t)niter_] / r0);
    else
      std::printf("Solver failed to converge after %5d iterations\n", params.maxiter);
  }
}

// a one-level handle for the smoothers and samplers (the level is swept, no
// coarse factorisation), a two-level one for the intergrid operator
static MultigridParameters level_params(int nlevel, double omega) {
  MultigridParameters p;
  p.nlevel = nlevel;
  p.omega = omega;
  p.npresmooth = p.npostsmooth = 0;
  p.coarse_solver = "SSOR";
  return p;
}

static void check_n(size_t a, size_t b, const char* what) {
  if (a != b) throw Error(MGMC_ERR_ARG, std::string(what) + ": vectors must hold one value per vertex");
}

HipSORSmoother::HipSORSmoother(std::shared_ptr<LinearOperator> linear_operator_, double omega, int nsmooth_,
                               Direction direction_, const DeviceOptions& o)
    : Smoother(std::move(linear_operator_)), nsmooth(nsmooth_), direction(direction_) {
  h = std::make_unique<Handle>(*linear_operator, level_params(1, omega), MGMC_COARSE_SSOR, o, 0);
}

void HipSORSmoother::apply(ConstVec b, Vec x) const {
  const size_t N = (size_t)linear_operator->get_lattice().Nvertex();
  check_n(b.n, N, "Smoother::apply");
  check_n(x.n, N, "Smoother::apply");
  check(mgmc_smoother_apply(h->get(), 0, MGMC_SOR, direction == Direction::forward, nsmooth, b.p, x.p),
        "mgmc_smoother_apply");
}

HipSSORSmoother::HipSSORSmoother(std::shared_ptr<LinearOperator> linear_operator_, double omega, int nsmooth_,
                                 const DeviceOptions& o)
    : Smoother(std::move(linear_operator_)), nsmooth(nsmooth_) {
  h = std::make_unique<Handle>(*linear_operator, level_params(1, omega), MGMC_COARSE_SSOR, o, 0);
}

void HipSSORSmoother::apply(ConstVec b, Vec x) const {
  const size_t N = (size_t)linear_operator->get_lattice().Nvertex();
  check_n(b.n, N, "Smoother::apply");
  check_n(x.n, N, "Smoother::apply");
  check(mgmc_smoother_apply(h->get(), 0, MGMC_SSOR, 1, nsmooth, b.p, x.p), "mgmc_smoother_apply");
}

HipSORSampler::HipSORSampler(std::shared_ptr<LinearOperator> linear_operator_, std::shared_ptr<RandomGenerator> rng_,
                             double omega_, int nsmooth_, Direction direction_, const DeviceOptions& o)
    : HipSORSampler(std::move(linear_operator_), std::move(rng_), omega_, nsmooth_, direction_, o, false) {}

HipSORSampler::HipSORSampler(std::shared_ptr<LinearOperator> linear_operator_, std::shared_ptr<RandomGenerator> rng_,
                             double omega_, int nsmooth_, Direction direction_, const DeviceOptions& o,
                             bool symmetric_)
    : Sampler(std::move(linear_operator_), std::move(rng_)), omega(omega_), nsmooth(nsmooth_),
      direction(direction_), options(o), symmetric(symmetric_) {
  if (!rng) throw Error(MGMC_ERR_ARG, "the sampler needs a random generator");
  key = draw_key(*rng);
  h = std::make_unique<Handle>(*linear_operator, level_params(1, omega), MGMC_COARSE_SSOR, options, key);
}

std::shared_ptr<Sampler> HipSORSampler::deep_copy(std::shared_ptr<RandomGenerator> rng_) {
  return std::make_shared<HipSORSampler>(linear_operator, std::move(rng_), omega, nsmooth, direction, options);
}

void HipSORSampler::apply(ConstVec f, Vec x) const {
  const size_t N = (size_t)linear_operator->get_lattice().Nvertex();
  check_n(f.n, N, "Sampler::apply");
  check_n(x.n, N, "Sampler::apply");
  check(mgmc_level_sample(h->get(), 0, symmetric ? MGMC_SSOR : MGMC_SOR, direction == Direction::forward, nsmooth,
                          f.p, x.p),
        "mgmc_level_sample");
}

HipSSORSampler::HipSSORSampler(std::shared_ptr<LinearOperator> linear_operator_,
                               std::shared_ptr<RandomGenerator> rng_, double omega_, int nsmooth_,
                               const DeviceOptions& o)
    : HipSORSampler(std::move(linear_operator_), std::move(rng_), omega_, nsmooth_, Direction::forward, o, true) {}

std::shared_ptr<Sampler> HipSSORSampler::deep_copy(std::shared_ptr<RandomGenerator> rng_) {
  return std::make_shared<HipSSORSampler>(linear_operator, std::move(rng_), omega, nsmooth, options);
}

HipIntergridOperator::HipIntergridOperator(std::shared_ptr<LinearOperator> fine_operator, const DeviceOptions& o) {
  h = std::make_unique<Handle>(*fine_operator, level_params(2, 1.0), MGMC_COARSE_SSOR, o, 0);
  check(mgmc_ndof(h->get(), 0, &n), "mgmc_ndof");
  check(mgmc_ndof(h->get(), 1, &nc), "mgmc_ndof");
}

void HipIntergridOperator::restrict(ConstVec x, Vec x_coarse) const {
  check_n(x.n, (size_t)n, "IntergridOperator::restrict");
  check_n(x_coarse.n, (size_t)nc, "IntergridOperator::restrict");
  check(mgmc_restrict(h->get(), 0, x.p, x_coarse.p), "mgmc_restrict");
}

void HipIntergridOperator::prolongate_add(double alpha, ConstVec x_coarse, Vec x) const {
  check_n(x.n, (size_t)n, "IntergridOperator::prolongate_add");
  check_n(x_coarse.n, (size_t)nc, "IntergridOperator::prolongate_add");
  check(mgmc_prolongate_add(h->get(), 0, alpha, x_coarse.p, x.p), "mgmc_prolongate_add");
}

}  // namespace mgmc_amd
