#include "driver_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <tuple>

namespace mgmc_host {

void die(const std::string& msg) {
  std::cout << "ERROR: " << msg << std::endl;
  std::exit(-1);
}

void check(int status, const char* what) {
  if (status != MGMC_OK) die(std::string(what) + ": " + mgmc_last_error());
}

// ---------------------------------------------------------------------------
// configuration groups (src/auxilliary/parameters.cc:52-350)
// ---------------------------------------------------------------------------

template <typename F>
static void guarded(const char* group, F f) {
  try {
    f();
  } catch (const ConfigError& e) {
    die(std::string("cannot read group '") + group + "' from the configuration: " + e.what());
  }
}

void GeneralParameters::parse(const Setting& root) {
  guarded("general", [&] {
    const Setting& g = root.lookup("general");
    dim = (int)g.lookup("dim").as_int();
    do_cholesky = g.lookup("do_cholesky").as_bool();
    do_ssor = g.lookup("do_ssor").as_bool();
    do_multigridmc = g.lookup("do_multigridmc").as_bool();
    measure_time_per_sample = g.lookup("measure_time_per_sample").as_bool();
    save_posterior_statistics = g.lookup("save_posterior_statistics").as_bool();
    measure_convergence = g.lookup("measure_convergence").as_bool();
    measure_mse = g.lookup("measure_mse").as_bool();
    operator_name = g.lookup("operator").as_string();
  });
  if (operator_name != "prior" && operator_name != "posterior") die("operator has to be 'prior' or 'posterior'");
  std::cout << "  dimension = " << dim << std::endl;
  std::cout << "  operator = " << operator_name << std::endl;
}

void LatticeParameters::parse(const Setting& root) {
  guarded("lattice", [&] {
    const Setting& l = root.lookup("lattice");
    nx = (int)l.lookup("nx").as_int();
    ny = (int)l.lookup("ny").as_int();
    nz = (int)l.lookup("nz").as_int();
  });
  std::cout << "  lattice size = " << nx << " x " << ny << " x " << nz << std::endl;
}

void CholeskyParameters::parse(const Setting& root) {
  guarded("cholesky", [&] { factorisation = root.lookup("cholesky.factorisation").as_string(); });
  if (factorisation != "sparse" && factorisation != "lowrank" && factorisation != "dense")
    die("Unknown Cholesky factorisation: '" + factorisation + "'");
  std::cout << "  Cholesky factorisation = " << factorisation << std::endl;
}

void SmootherParameters::parse(const Setting& root) {
  guarded("smoother", [&] {
    omega = root.lookup("smoother.omega").as_double();
    nsmooth = (int)root.lookup("smoother.nsmooth").as_int();
  });
  std::cout << "  smoother/Gibbs sampler " << std::endl;
  std::cout << "    number of smoothing steps = " << nsmooth << std::endl;
  std::cout << "    overrelaxation factor = " << omega << std::endl;
}

void IterativeSolverParameters::parse(const Setting& root) {
  guarded("iterative_solver", [&] {
    const Setting& s = root.lookup("iterative_solver");
    rtol = s.lookup("rtol").as_double();
    atol = s.lookup("atol").as_double();
    maxiter = (int)s.lookup("maxiter").as_int();
    verbose = (int)s.lookup("verbose").as_int();
  });
  std::cout << "  iterative solver " << std::endl;
  std::cout << "    relative tolerance = " << rtol << std::endl;
  std::cout << "    absolute tolerance = " << atol << std::endl;
  std::cout << "    maxiter = " << maxiter << std::endl;
}

void MultigridParameters::parse(const Setting& root) {
  guarded("multigrid", [&] {
    const Setting& m = root.lookup("multigrid");
    nlevel = (int)m.lookup("nlevel").as_int();
    smoother = m.lookup("smoother").as_string();
    coarse_solver = m.lookup("coarse_solver").as_string();
    npresmooth = (int)m.lookup("npresmooth").as_int();
    npostsmooth = (int)m.lookup("npostsmooth").as_int();
    ncoarsesmooth = (int)m.lookup("ncoarsesmooth").as_int();
    cycle = (int)m.lookup("cycle").as_int();
    coarse_scaling = m.lookup("coarse_scaling").as_double();
    omega = m.lookup("omega").as_double();
    verbose = (int)m.lookup("verbose").as_int();
  });
  if (smoother != "SOR" && smoother != "SSOR") die("invalid multigrid smoother : '" + smoother + "'");
  if (coarse_solver != "SSOR" && coarse_solver != "Cholesky")
    die("invalid multigrid coarse solver : '" + coarse_solver + "'");
  std::cout << "  multigrid" << std::endl;
  std::cout << "    levels = " << nlevel << std::endl;
  std::cout << "    npresmooth = " << npresmooth << std::endl;
  std::cout << "    npostsmooth = " << npostsmooth << std::endl;
  std::cout << "    ncoarsesmooth = " << ncoarsesmooth << std::endl;
  std::cout << "    smoother = " << smoother << std::endl;
  std::cout << "    coarse solver = " << coarse_solver << std::endl;
  std::cout << "    overrelaxation factor = " << omega << std::endl;
  std::cout << "    cycle = " << cycle << " " << (cycle == 1 ? "( V-cycle )" : cycle == 2 ? "( W-cycle )" : "")
            << std::endl;
  std::cout << "    coarse_scaling = " << coarse_scaling << std::endl;
}

void SamplingParameters::parse(const Setting& root) {
  guarded("sampling", [&] {
    const Setting& s = root.lookup("sampling");
    nsamples = (int)s.lookup("timeseries.nsamples").as_int();
    nwarmup = (int)s.lookup("timeseries.nwarmup").as_int();
    nstepsconvergence = (int)s.lookup("convergence.nsteps").as_int();
    nsamplesconvergence = (int)s.lookup("convergence.nsamples").as_int();
    nthreadsconvergence = (int)s.lookup("convergence.nthreads").as_int();
    nstepsmse = (int)s.lookup("mse.nsteps").as_int();
    nsamplesmse = (int)s.lookup("mse.nsamples").as_int();
  });
  std::cout << "  timeseries: number of samples         = " << nsamples << std::endl;
  std::cout << "  timeseries: number of warmup samples  = " << nwarmup << std::endl;
  std::cout << "  convergence test: number of steps     = " << nstepsconvergence << std::endl;
  std::cout << "  convergence test: number of samples   = " << nsamplesconvergence << std::endl;
  std::cout << "  MSE test: number of steps             = " << nstepsmse << std::endl;
  std::cout << "  MSE test: number of samples           = " << nsamplesmse << std::endl;
}

void PriorParameters::parse(const Setting& root) {
  guarded("prior", [&] {
    pde_model = root.lookup("prior.pdemodel").as_string();
    correlationlength_model = root.lookup("prior.correlationlengthmodel").as_string();
  });
  if (pde_model != "shiftedlaplace_fem" && pde_model != "shiftedlaplace_fd" &&
      pde_model != "squared_shiftedlaplace_fd")
    die("Unknown prior: '" + pde_model + "'");
  if (correlationlength_model != "constant" && correlationlength_model != "periodic")
    die("Invalid correlation length model '" + correlationlength_model + "'.");
  if (correlationlength_model == "constant")
    guarded("constantcorrelationlengthmodel",
            [&] { Lambda = root.lookup("constantcorrelationlengthmodel.Lambda").as_double(); });
  else
    guarded("periodiccorrelationlengthmodel", [&] {
      Lambda_min = root.lookup("periodiccorrelationlengthmodel.Lambda_min").as_double();
      Lambda_max = root.lookup("periodiccorrelationlengthmodel.Lambda_max").as_double();
    });
  std::cout << "  prior" << std::endl;
  std::cout << "    PDEmodel = " << pde_model << std::endl;
  std::cout << "    correlationlengthmodel = " << correlationlength_model << std::endl;
  if (correlationlength_model == "constant") std::cout << "    Lambda = " << Lambda << std::endl;
}

void MeasurementParameters::parse(const Setting& root) {
  guarded("measurements", [&] {
    const Setting& m = root.lookup("measurements");
    radius = m.lookup("radius").as_double();
    variance_scaling = m.lookup("variance_scaling").as_double();
    measure_global = m.lookup("measure_global").as_bool();
    variance_global = m.lookup("variance_global").as_double();
    mean_global = m.lookup("mean_global").as_double();
    filename = m.lookup("filename").as_string();
    sample_location = m.lookup("sample_location").as_doubles();
  });
  Setting data;
  try {
    data = read_config_file(filename);
  } catch (const ConfigError& e) {
    die("opening configuration file with measurements: '" + filename + "' (" + e.what() + ")");
  }
  try {
    dim = (int)data.lookup("dim").as_int();
    n = (int)data.lookup("n").as_int();
    std::vector<double> loc = data.lookup("measurement_locations").as_doubles();
    std::vector<double> mu = data.lookup("mean").as_doubles();
    std::vector<double> var = data.lookup("variance").as_doubles();
    if ((int)loc.size() < n * dim || (int)mu.size() < n || (int)var.size() < n)
      throw ConfigError("fewer entries than n measurements");
    locations.assign(loc.begin(), loc.begin() + (size_t)n * dim);
    mean.assign(mu.begin(), mu.begin() + n);
    variance.assign(var.begin(), var.begin() + n);
  } catch (const ConfigError& e) {
    die("reading configuration file with measurements: '" + filename + "' (" + e.what() + ")");
  }
  std::cout << "  measurements " << std::endl;
  std::cout << "    file with measurements = '" << filename << "'" << std::endl;
  std::cout << "    dimension of measurement locations = " << dim << std::endl;
  std::cout << "    number of measurement points = " << n << std::endl;
  std::cout << "    radius of individual measurements = " << radius << std::endl;
  std::cout << "    variance scaling = " << variance_scaling << std::endl;
  std::cout << "    measure global average across domain? " << (measure_global ? "yes" : "no") << std::endl;
  std::cout << std::endl;
}

void GpuParameters::parse(const Setting& root) {
  if (!root.exists("gpu")) return;
  guarded("gpu", [&] {
    const Setting& g = root.lookup("gpu");
    if (g.exists("nchains")) nchains = (int)g.lookup("nchains").as_int();
    if (g.exists("convergence_chains")) convergence_chains = (int)g.lookup("convergence_chains").as_int();
    if (g.exists("precision")) precision = g.lookup("precision").as_string();
    if (g.exists("seed")) seed = (uint64_t)g.lookup("seed").as_int();
    if (g.exists("ordering")) ordering = g.lookup("ordering").as_string();
  });
  if (precision != "fp64" && precision != "fp32") die("gpu.precision has to be 'fp64' or 'fp32'");
  if (ordering != "multicolour" && ordering != "lexicographic")
    die("gpu.ordering has to be 'multicolour' or 'lexicographic'");
  if (nchains < 1 || convergence_chains < 1) die("gpu.nchains and gpu.convergence_chains must be >= 1");
  std::cout << "  gpu: chains = " << nchains << ", convergence chains per batch = " << convergence_chains
            << ", precision = " << precision << ", ordering = " << ordering << ", seed = " << seed << std::endl;
}

// ---------------------------------------------------------------------------
// problem set-up (src/driver_mgmc.cc:91-161)
// ---------------------------------------------------------------------------

mgmc_problem Problem::view() const {
  mgmc_problem p{};
  p.dim = dim;
  for (int d = 0; d < 3; ++d) p.n[d] = n[d];
  p.kappa_sq = kappa_sq;
  if (!kappa_field.empty()) p.kappa_sq_field = kappa_field.data();
  if (fem) {
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

static void euclid(const Problem& P, int64_t ell, int* e) {
  for (int d = 0; d < P.dim; ++d) {
    e[d] = (int)(ell % (P.n[d] - 1)) + 1;
    ell /= (P.n[d] - 1);
  }
}

static int64_t linear(const Problem& P, const int* e) {
  int64_t ell = 0, stride = 1;
  for (int d = 0; d < P.dim; ++d) {
    ell += (int64_t)(e[d] - 1) * stride;
    stride *= P.n[d] - 1;
  }
  return ell;
}

Observation measurement_vector(const Problem& P, const std::vector<double>& x0, double radius) {
  if ((int)x0.size() != P.dim) die("measurement location has the wrong dimension");
  Observation o;
  int64_t nnz = 0;
  int64_t cap = 1024;
  for (;;) {
    o.idx.resize(cap);
    o.val.resize(cap);
    int st = mgmc_measurement_vector(P.dim, P.n, x0.data(), radius, cap, o.idx.data(), o.val.data(), &nnz);
    if (st == MGMC_OK) break;
    if (nnz <= cap) check(st, "measurement vector");
    cap = nnz;
  }
  o.idx.resize(nnz);
  o.val.resize(nnz);
  return o;
}

Problem build_problem(const GeneralParameters& gp, const LatticeParameters& lp, const PriorParameters& pp,
                      const MeasurementParameters& mp) {
  Problem P;
  P.dim = gp.dim;
  if (P.dim != 2 && P.dim != 3) die("Invalid dimension : " + std::to_string(P.dim));
  P.n[0] = lp.nx;
  P.n[1] = lp.ny;
  P.n[2] = P.dim == 3 ? lp.nz : 1;
  P.N = 1;
  for (int d = 0; d < P.dim; ++d) P.N *= P.n[d] - 1;
  if (pp.correlationlength_model == "constant") {
    P.kappa_sq = 1.0 / (pp.Lambda * pp.Lambda);
  } else {
    // PeriodicCorrelationLengthModel (src/linear_operator/correlationlength_model.hh:83-105):
    // Lambda(x) = Lambda_1 + Lambda_2 prod_d cos(pi x_d), kappa^2 = 1 / Lambda^2 at every vertex
    if (pp.pde_model != "shiftedlaplace_fd")
      die("the periodic correlation length model is supported with the finite-difference prior only");
    const double L1 = 0.5 * (pp.Lambda_max + pp.Lambda_min), L2 = 0.5 * (pp.Lambda_max - pp.Lambda_min);
    P.kappa_field.resize((size_t)P.N);
    for (int64_t r = 0; r < P.N; ++r) {
      int e[3] = {1, 1, 1};
      euclid(P, r, e);
      double lam = L2;
      for (int d = 0; d < P.dim; ++d) lam *= std::cos(M_PI * ((double)e[d] * (1. / (double)P.n[d])));
      lam += L1;
      P.kappa_field[(size_t)r] = 1. / (lam * lam);
    }
  }
  if (pp.pde_model == "squared_shiftedlaplace_fd")
    die("the squared shifted-Laplace prior is not supported by the GPU library");
  if (pp.pde_model == "shiftedlaplace_fem") {
    double st[27];
    check(mgmc_fem_stencil(P.dim, P.n, P.kappa_sq, st), "FEM stencil");
    P.fem = true;
    P.A_rowptr.assign(1, 0);
    const int zr = P.dim == 3 ? 1 : 0;
    for (int64_t r = 0; r < P.N; ++r) {
      int e[3] = {1, 1, 1};
      euclid(P, r, e);
      // columns in increasing order: z slowest, x fastest
      for (int dz = -zr; dz <= zr; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int b[3] = {e[0] + dx, e[1] + dy, e[2] + dz};
            bool ok = true;
            for (int d = 0; d < P.dim; ++d) ok = ok && b[d] >= 1 && b[d] <= P.n[d] - 1;
            if (!ok) continue;
            P.A_col.push_back(linear(P, b));
            P.A_val.push_back(P.dim == 3 ? st[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] : st[(dy + 1) * 3 + (dx + 1)]);
          }
      P.A_rowptr.push_back((int64_t)P.A_col.size());
    }
  }
  if (gp.operator_name == "posterior") {
    if (mp.measure_global) die("the global-average measurement (measure_global) is not supported by the GPU library");
    if (mp.dim != P.dim) die("dimension of measurement locations differs from problem dimension");
    P.m = mp.n;
    std::vector<std::tuple<int64_t, int64_t, double>> trip;
    for (int k = 0; k < mp.n; ++k) {
      std::vector<double> x0(mp.locations.begin() + (size_t)k * P.dim, mp.locations.begin() + (size_t)(k + 1) * P.dim);
      Observation col = measurement_vector(P, x0, mp.radius);
      for (size_t e = 0; e < col.idx.size(); ++e) trip.emplace_back(col.idx[e], k, col.val[e]);
      P.sigma.push_back(mp.variance_scaling * mp.variance[k]);
      P.y.push_back(mp.mean[k]);
    }
    std::sort(trip.begin(), trip.end());
    P.B_rowptr.assign((size_t)P.N + 1, 0);
    for (const auto& t : trip) {
      P.B_rowptr[(size_t)std::get<0>(t) + 1]++;
      P.B_col.push_back(std::get<1>(t));
      P.B_val.push_back(std::get<2>(t));
    }
    for (int64_t r = 0; r < P.N; ++r) P.B_rowptr[r + 1] += P.B_rowptr[r];
  }
  return P;
}

// ---------------------------------------------------------------------------
// handles and samplers
// ---------------------------------------------------------------------------

Handle::Handle(const Problem& P, const mgmc_params& par) : nchains(par.nchains) {
  mgmc_problem p = P.view();
  check(mgmc_create(&p, &par, &h), "mgmc_create");
}

Handle::Handle(const SamplerSpec& s, int nc) : nchains(nc) {
  mgmc_problem p = s.P->view();
  mgmc_params par = s.par;
  par.nchains = nc;
  check(mgmc_create(&p, &par, &h), "mgmc_create");
}

Handle::~Handle() {
  if (h) mgmc_destroy(h);
}

static mgmc_params base_params(const GpuParameters& gpu) {
  mgmc_params par{};
  par.nlevel = 1;
  par.smoother = MGMC_SSOR;
  par.npresmooth = 1;
  par.npostsmooth = 1;
  par.ncoarsesmooth = 1;
  par.omega = 1.0;
  par.cycle = 1;
  par.coarse_scaling = 1.0;
  par.coarse_solver = MGMC_COARSE_CHOLESKY;
  par.precision = gpu.precision == "fp32" ? MGMC_FP32 : MGMC_FP64;
  par.nchains = 1;
  par.seed = gpu.seed;
  par.stream = nullptr;
  return par;
}

SamplerSpec multigridmc_sampler(const Problem& P, const MultigridParameters& mg, const GpuParameters& gpu) {
  mgmc_params par = base_params(gpu);
  par.nlevel = mg.nlevel;
  par.smoother = mg.smoother == "SSOR" ? MGMC_SSOR : MGMC_SOR;
  par.npresmooth = mg.npresmooth;
  par.npostsmooth = mg.npostsmooth;
  par.ncoarsesmooth = mg.ncoarsesmooth;
  par.omega = mg.omega;
  par.cycle = mg.cycle;
  par.coarse_scaling = mg.coarse_scaling;
  par.coarse_solver = mg.coarse_solver == "Cholesky" ? MGMC_COARSE_CHOLESKY : MGMC_COARSE_SSOR;
  return SamplerSpec{&P, par};
}

SamplerSpec ssor_sampler(const Problem& P, const SmootherParameters& sm, const GpuParameters& gpu) {
  // SSORSampler (src/sampler/ssor_sampler.cc:9-15): a one-level hierarchy
  // whose level is sampled by nsmooth SSOR sweeps
  mgmc_params par = base_params(gpu);
  par.npresmooth = 0;
  par.npostsmooth = 0;
  par.ncoarsesmooth = sm.nsmooth;
  par.omega = sm.omega;
  par.coarse_solver = MGMC_COARSE_SSOR;
  return SamplerSpec{&P, par};
}

SamplerSpec cholesky_sampler(const Problem& P, const CholeskyParameters& ch, const GpuParameters& gpu) {
  SamplerSpec s{&P, base_params(gpu)};
  s.cholesky = ch.factorisation == "lowrank" ? MGMC_FACTOR_LOWRANK
               : ch.factorisation == "dense" ? MGMC_FACTOR_DENSE
                                             : MGMC_FACTOR_SPARSE;
  mgmc_problem p = P.view();
  mgmc_cholesky* c = nullptr;
  auto t0 = std::chrono::high_resolution_clock::now();
  check(mgmc_cholesky_create(&p, s.cholesky, 1, gpu.seed, 0, &c), "mgmc_cholesky_create");
  auto t1 = std::chrono::high_resolution_clock::now();
  s.factor = std::shared_ptr<mgmc_cholesky>(c, [](mgmc_cholesky* q) { mgmc_cholesky_destroy(q); });
  int64_t nnz = 0;
  check(mgmc_cholesky_info(c, nullptr, &nnz, nullptr, nullptr), "mgmc_cholesky_info");
  std::cout << std::endl;
  std::cout << "  matrix size = " << P.N << " x " << P.N << ", nnz(L) = " << nnz << std::endl;
  std::cout << "time for Cholesky factorisation = " << std::chrono::duration<double>(t1 - t0).count() << " s"
            << std::endl;
  return s;
}

namespace {

struct GpuChains : Chains {
  Handle s;
  const Problem& P;
  GpuChains(const SamplerSpec& spec, int K) : s(spec, K), P(*spec.P) { nchains = K; }
  void prepare(const Exact& ex, const Observation& b) override {
    check(mgmc_set_rhs(s.h, ex.f.data(), 0), "mgmc_set_rhs");
    std::vector<double> x0((size_t)P.N * nchains, 0.0);
    check(mgmc_set_state(s.h, x0.data()), "mgmc_set_state");
    int64_t rp[2] = {0, (int64_t)b.idx.size()};
    check(mgmc_set_observables(s.h, 1, rp, b.idx.data(), b.val.data()), "mgmc_set_observables");
  }
  void run(int nsteps, double* obs) override {
    if (nsteps > 0) check(mgmc_run(s.h, nsteps, obs, 1), "mgmc_run");
  }
  void moments(int nsamples, double* mean, double* var) override {
    check(mgmc_moments(s.h, nsamples, mean, var), "mgmc_moments");
  }
};

struct CholeskyChains : Chains {
  std::shared_ptr<mgmc_cholesky> factor;
  mgmc_cholesky* c = nullptr;
  const Problem& P;
  Observation b;
  std::vector<double> x;
  CholeskyChains(const SamplerSpec& spec, int K) : factor(spec.factor), c(spec.factor.get()), P(*spec.P) {
    nchains = K;
    x.resize((size_t)P.N * K);
  }
  void prepare(const Exact& ex, const Observation& obs) override {
    check(mgmc_cholesky_set_nchains(c, nchains), "mgmc_cholesky_set_nchains");
    check(mgmc_cholesky_fix_rhs(c, ex.f.data()), "mgmc_cholesky_fix_rhs");
    b = obs;
  }
  void run(int nsteps, double* obs) override {
    for (int k = 0; k < nsteps; ++k) {
      check(mgmc_cholesky_sample(c, nullptr, 0, nullptr, x.data()), "mgmc_cholesky_sample");
      if (!obs) continue;
      for (int ch = 0; ch < nchains; ++ch) {
        double z = 0.0;
        for (size_t e = 0; e < b.idx.size(); ++e) z += b.val[e] * x[(size_t)ch * P.N + b.idx[e]];
        obs[(size_t)k * nchains + ch] = z;
      }
    }
  }
  void moments(int nsamples, double* mean, double* var) override {
    std::vector<double> s1((size_t)P.N, 0.0), s2((size_t)P.N, 0.0);
    for (int k = 0; k < nsamples; ++k) {
      check(mgmc_cholesky_sample(c, nullptr, 0, nullptr, x.data()), "mgmc_cholesky_sample");
      for (int ch = 0; ch < nchains; ++ch)
        for (int64_t i = 0; i < P.N; ++i) {
          const double v = x[(size_t)ch * P.N + i];
          s1[i] += v;
          s2[i] += v * v;
        }
    }
    const double n = (double)nsamples * nchains;
    for (int64_t i = 0; i < P.N; ++i) {
      const double m = s1[i] / n;
      if (mean) mean[i] = m;
      if (var) var[i] = s2[i] / n - m * m;
    }
  }
};

}  // namespace

std::unique_ptr<Chains> make_chains(const SamplerSpec& s, int nchains) {
  if (s.cholesky >= 0) return std::make_unique<CholeskyChains>(s, nchains);
  return std::make_unique<GpuChains>(s, nchains);
}

// ---------------------------------------------------------------------------
// exact statistics (linear_operator.hh:131-189) by FP64 multigrid solves
// ---------------------------------------------------------------------------

static void solve(mgmc_handle* h, const std::vector<double>& b, std::vector<double>& x) {
  std::vector<double> hist(200);
  int niter = -1;
  x.assign(b.size(), 0.0);
  int st = mg_solve(h, b.data(), x.data(), 1e-13, 1e300, 200, hist.data(), &niter);
  if (st != MGMC_OK && st != MGMC_ERR_NOCONV) check(st, "mg_solve");
}

Exact exact_statistics(const Problem& P, const MultigridParameters& mg, const Observation& b) {
  GpuParameters g;
  SamplerSpec spec = multigridmc_sampler(P, mg, g);
  spec.par.precision = MGMC_FP64;
  spec.par.coarse_solver = MGMC_COARSE_CHOLESKY;  // the preconditioner's coarse solve is Cholesky
  Handle h(spec, 1);
  Exact ex;
  ex.x_post.assign((size_t)P.N, 0.0);
  if (P.m > 0) {
    std::vector<double> rhs((size_t)P.N, 0.0);
    for (int64_t r = 0; r < P.N; ++r)
      for (int64_t e = P.B_rowptr[r]; e < P.B_rowptr[r + 1]; ++e) rhs[r] += P.B_val[e] * (P.y[P.B_col[e]] / P.sigma[P.B_col[e]]);
    solve(h.h, rhs, ex.x_post);
  }
  ex.f.assign((size_t)P.N, 0.0);
  check(mgmc_apply_operator(h.h, ex.x_post.data(), ex.f.data()), "mgmc_apply_operator");
  std::vector<double> bd((size_t)P.N, 0.0), w;
  for (size_t e = 0; e < b.idx.size(); ++e) bd[b.idx[e]] += b.val[e];
  solve(h.h, bd, w);
  double mean = 0.0, var = 0.0;
  for (int64_t i = 0; i < P.N; ++i) {
    mean += bd[i] * ex.x_post[i];
    var += bd[i] * w[i];
  }
  ex.mean = mean;
  ex.variance = var;
  return ex;
}

// ---------------------------------------------------------------------------
// harnesses (src/measurements/*.cc)
// ---------------------------------------------------------------------------

void measure_sampling_time(const SamplerSpec& spec, const Exact& ex, const Observation& b, const SamplingParameters& sp,
                           const GpuParameters& gpu, const std::string& label, const std::string& filename) {
  std::unique_ptr<Chains> s = make_chains(spec, gpu.nchains);
  s->prepare(ex, b);
  s->run(sp.nwarmup, nullptr);
  const int K = s->nchains;
  std::vector<double> data((size_t)sp.nsamples * K);
  auto t0 = std::chrono::high_resolution_clock::now();
  s->run(sp.nsamples, data.data());
  auto t1 = std::chrono::high_resolution_clock::now();
  double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / ((double)sp.nsamples * K);
  printf("  %12s time per sample = %12.4f ms\n", label.c_str(), ms);
  std::ofstream out(filename);
  for (int k = 0; k < sp.nsamples; ++k) {
    for (int c = 0; c < K; ++c) out << (c ? " " : "") << data[(size_t)k * K + c];
    out << std::endl;
  }
  double x_avg = 0.0, xsq_avg = 0.0;
  for (size_t k = 0; k < data.size(); ++k) {
    x_avg += (data[k] - x_avg) / (k + 1.0);
    xsq_avg += (data[k] * data[k] - xsq_avg) / (k + 1.0);
  }
  double variance = xsq_avg - x_avg * x_avg;
  double x_error = std::sqrt(variance / data.size());
  printf("  %12s mean     = %12.4e +/- %12.4e [ignoring IACT]\n", label.c_str(), x_avg, x_error);
  printf("  %12s mean     = %12.4e\n", "exact", ex.mean);
  printf("  %12s variance = %12.4e\n", label.c_str(), variance);
  printf("  %12s variance = %12.4e\n\n", "exact", ex.variance);
}

// z at steps 0..nsteps of `nsamples` chains started from x = 0, in batches
static std::vector<double> series(const SamplerSpec& spec, const Exact& ex, const Observation& b, int nsamples,
                                  int nsteps, int batch) {
  std::vector<double> z((size_t)nsamples * (nsteps + 1), 0.0);
  int done = 0;
  while (done < nsamples) {
    int K = std::min(batch, nsamples - done);
    std::unique_ptr<Chains> s = make_chains(spec, K);
    s->prepare(ex, b);
    std::vector<double> obs((size_t)std::max(nsteps, 1) * K);
    s->run(nsteps, obs.data());
    for (int c = 0; c < K; ++c)
      for (int j = 1; j <= nsteps; ++j) z[(size_t)(done + c) * (nsteps + 1) + j] = obs[(size_t)(j - 1) * K + c];
    done += K;
  }
  return z;
}

void measure_convergence(const SamplerSpec& spec, const Exact& ex, const Observation& b, const SamplingParameters& sp,
                         const GpuParameters& gpu, const std::string& filename) {
  const int n = sp.nsamplesconvergence, ns = sp.nstepsconvergence;
  std::vector<double> z = series(spec, ex, b, n, ns, gpu.convergence_chains);
  std::vector<double> m1(ns + 1, 0.0), m2(ns + 1, 0.0), m3(ns + 1, 0.0), m4(ns + 1, 0.0);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j <= ns; ++j) {
      double v = z[(size_t)k * (ns + 1) + j];
      m1[j] += (v - m1[j]) / (k + 1.0);
      m2[j] += (v * v - m2[j]) / (k + 1.0);
      m3[j] += (v * v * v - m3[j]) / (k + 1.0);
      m4[j] += (v * v * v * v - m4[j]) / (k + 1.0);
    }
  std::vector<double> dm, dv, em, ev;
  for (int j = 0; j <= ns; ++j) {
    dm.push_back(std::fabs(m1[j] - ex.mean));
    dv.push_back(std::fabs(m2[j] - m1[j] * m1[j] - ex.variance));
    double sigma_sq = n / (n - 1.) * (m2[j] - m1[j] * m1[j]);
    double mu4 = m4[j] - 4 * m1[j] * m3[j] + 6 * std::pow(m1[j], 2) * m2[j] - 3 * std::pow(m1[j], 4);
    em.push_back(std::sqrt(sigma_sq / n));
    ev.push_back(std::sqrt((mu4 - (n - 3.) / (n - 1.) * sigma_sq * sigma_sq) / n));
  }
  std::ofstream out(filename);
  char buf[256];
  for (int q = 0; q < 2; ++q) {
    const std::vector<double>& d = q == 0 ? dm : dv;
    const std::vector<double>& e = q == 0 ? em : ev;
    const char* label = q == 0 ? "mean" : "variance";
    out << (q == 0 ? "**** q_k = |E[z^k] - E[z]| **** " : "**** q_k = |Var[z^k] - Var[z]| **** ") << std::endl;
    snprintf(buf, sizeof(buf), "  %12s   %3s : %12s %35s %35s\n", "", "k", "q_k", "q_k/q_0", "q_k/q_{k-1}");
    out << buf;
    for (int j = 0; j <= ns; ++j) {
      snprintf(buf, sizeof(buf), "  %12s   %3d : %12.8f +/- %12.8f       %12.8f +/- %12.8f      ", label, j, d[j], e[j],
               d[j] / d[0], e[j] / d[0]);
      out << buf;
      if (j > 0) {
        double rel = d[j] / d[j - 1] * std::sqrt(std::pow(e[j] / d[j], 2) + std::pow(e[j - 1] / d[j - 1], 2));
        snprintf(buf, sizeof(buf), " %12.8f +/- %12.8f \n", d[j] / d[j - 1], rel);
      } else {
        snprintf(buf, sizeof(buf), " %12s\n", "---");
      }
      out << buf;
    }
    out << std::endl;
  }
}

void measure_mean_squared_error(const SamplerSpec& spec, const Exact& ex, const Observation& b,
                                const SamplingParameters& sp, const GpuParameters& gpu, const std::string& filename) {
  std::cout << "  Measuring MSE" << std::endl;
  const int n = sp.nsamplesmse, ns = sp.nstepsmse;
  if (ns < 1) return;
  std::vector<double> z = series(spec, ex, b, n, ns - 1, gpu.convergence_chains);
  std::vector<double> x_avg(ns, 0.0), x2_avg(ns, 0.0);
  for (int k = 0; k < n; ++k) {
    double run = 0.0;
    for (int j = 0; j < ns; ++j) {
      run += (z[(size_t)k * ns + j] - run) / (j + 1);
      double mse = (run - ex.mean) * (run - ex.mean);
      x_avg[j] += (mse - x_avg[j]) / (k + 1.0);
      x2_avg[j] += (mse * mse - x2_avg[j]) / (k + 1.0);
    }
  }
  std::ofstream out(filename);
  char buf[256];
  for (int j = 0; j < ns; ++j) {
    double err = 1. / (n - 1.) * (x2_avg[j] - x_avg[j] * x_avg[j]);
    snprintf(buf, sizeof(buf), " %5d : %12.6e +/- %12.6e\n", j, x_avg[j], err);
    out << buf;
  }
}

void posterior_statistics(const SamplerSpec& spec, const Exact& ex, const SamplingParameters& sp,
                          const MeasurementParameters& mp, const GpuParameters& gpu) {
  const Problem& P = *spec.P;
  std::unique_ptr<Chains> s = make_chains(spec, gpu.nchains);
  Observation none;
  s->prepare(ex, none);
  s->run(sp.nwarmup, nullptr);
  std::vector<double> mean((size_t)P.N), var((size_t)P.N);
  s->moments(std::max(sp.nsamples, 1), mean.data(), var.data());
  write_vtk("posterior.vtk", P, {{"mean", &mean}, {"variance", &var}, {"mean_exact", &ex.x_post}});
  if (P.dim == 2) write_vtk_circle(mp.sample_location, mp.radius, "sample_location.vtk");
}

// ---------------------------------------------------------------------------
// VTK output (vtk_writer2d.cc, vtk_writer3d.cc)
// ---------------------------------------------------------------------------

void write_vtk(const std::string& filename, const Problem& P,
               const std::vector<std::pair<std::string, const std::vector<double>*>>& states) {
  const int nx = P.n[0], ny = P.n[1], nz = P.dim == 3 ? P.n[2] : 0;
  std::ofstream out(filename);
  out << "# vtk DataFile Version 2.0" << std::endl;
  out << "Sample state" << std::endl;
  out << "ASCII" << std::endl;
  out << "DATASET STRUCTURED_POINTS" << std::endl;
  if (P.dim == 3) {
    out << "DIMENSIONS " << nx + 1 << " " << ny + 1 << " " << nz + 1 << std::endl;
    out << "ORIGIN -0.5 -0.5 -5.0" << std::endl;
    out << "SPACING " << 1. / nx << " " << 1. / ny << " " << 1. / nz << std::endl;
    out << std::endl;
    out << "POINT_DATA " << (nx + 1) * (ny + 1) * (nz + 1) << std::endl;
  } else {
    out << "DIMENSIONS " << nx + 1 << " " << ny + 1 << " 1 " << std::endl;
    out << "ORIGIN -0.5 -0.5 0.0" << std::endl;
    out << "SPACING " << 1. / nx << " " << 1. / ny << " 0" << std::endl;
    out << std::endl;
    out << "POINT_DATA " << (nx + 1) * (ny + 1) << std::endl;
  }
  for (const auto& st : states) {
    out << "SCALARS " << st.first << " double 1" << std::endl;
    out << "LOOKUP_TABLE default" << std::endl;
    const std::vector<double>& phi = *st.second;
    for (int k = 0; k <= nz; ++k)
      for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
          double data = 0.0;
          bool interior = i > 0 && i < nx && j > 0 && j < ny && (P.dim == 2 || (k > 0 && k < nz));
          if (interior) {
            int e[3] = {i, j, k};
            data = phi[(size_t)linear(P, e)];
            if (std::fabs(data) < 1.0E-20) data = 0.0;
          }
          out << data << std::endl;
        }
  }
}

void write_vtk_circle(const std::vector<double>& centre, double radius, const std::string& filename) {
  std::ofstream out(filename);
  out << "# vtk DataFile Version 2.0" << std::endl;
  out << "Sample state" << std::endl;
  out << "ASCII" << std::endl;
  out << "DATASET POLYDATA" << std::endl;
  out << std::endl;
  const int npoints = 100;
  const double z_offset = 1.E-6;
  out << "POINTS " << npoints << " double" << std::endl;
  for (int j = 0; j < npoints; ++j) {
    double x = centre[0] + radius * std::cos(2 * M_PI * j / (1.0 * npoints)) - 0.5;
    double y = centre[1] + radius * std::sin(2 * M_PI * j / (1.0 * npoints)) - 0.5;
    out << x << " " << y << " " << z_offset << std::endl;
  }
  out << "POLYGONS 1 " << (npoints + 1) << std::endl;
  out << npoints;
  for (int j = 0; j < npoints; ++j) out << " " << j;
  out << std::endl;
}

}  // namespace mgmc_host
