// Runs libmgmc_amd's kernels on the CPU (hip_emu.h) under the sanitizers:
// sampling cycles (fused red-black sweeps with the folded coarse correction,
// box-level plane sweeps, low-rank kernels, the coarsest-level sample) and
// multigrid solves, in FP64 and FP32, multicolour and lexicographic order,
// with constant and per-vertex stencils, on lattices whose x extent is and is
// not a multiple of 8. A solve must converge; the samples must be finite.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/mgmc_amd.h"

namespace {

int failures = 0;

void check(int code, const char* what) {
  if (code != MGMC_OK) {
    std::printf("FAIL %s: %d %s\n", what, code, mgmc_last_error());
    std::exit(1);
  }
}

struct Case {
  const char* name;
  int dim, n[3], m, nlevel, precision, ordering, K;
  bool variable;
};

// The problem of a case: its measurement CSR, variances and diffusion
// field, kept alive while handles are created from it.
struct ProblemData {
  int64_t N = 1;
  std::vector<int64_t> rp, col;
  std::vector<double> bval, sigma, diffusion;
  mgmc_problem p{};
  mgmc_params par{};
};

void build(const Case& c, ProblemData& d) {
  int64_t N = 1;
  for (int dd = 0; dd < c.dim; ++dd) N *= c.n[dd] - 1;
  d.N = N;
  // measurements: balls of radius 0.2 at seeded centres
  std::vector<std::vector<std::pair<int64_t, double>>> rows((size_t)N);
  for (int k = 0; k < c.m; ++k) {
    double x0[3] = {0.3 + 0.17 * k, 0.6 - 0.11 * k, 0.45};
    std::vector<int64_t> idx((size_t)N);
    std::vector<double> val((size_t)N);
    int64_t nnz = 0;
    check(mgmc_measurement_vector(c.dim, c.n, x0, 0.2, N, idx.data(), val.data(), &nnz), "measurement_vector");
    for (int64_t e = 0; e < nnz; ++e) rows[(size_t)idx[e]].push_back({k, val[e]});
    d.sigma.push_back(1e-3 * (1.0 + k));
  }
  d.rp.assign(1, 0);
  for (auto& r : rows) {
    for (auto& e : r) {
      d.col.push_back(e.first);
      d.bval.push_back(e.second);
    }
    d.rp.push_back((int64_t)d.col.size());
  }
  int64_t ncell = 1;
  for (int dd = 0; dd < c.dim; ++dd) ncell *= c.n[dd];
  d.diffusion.resize((size_t)ncell);
  for (int64_t i = 0; i < ncell; ++i) d.diffusion[(size_t)i] = 1.0 + 9.0 * (double)((i * 7919) % 13) / 12.0;
  mgmc_problem& p = d.p;
  p.dim = c.dim;
  for (int dd = 0; dd < 3; ++dd) p.n[dd] = c.n[dd];
  p.kappa_sq = 10.0;
  p.m = c.m;
  if (c.m > 0) {
    p.B_rowptr = d.rp.data();
    p.B_col = d.col.data();
    p.B_val = d.bval.data();
    p.sigma = d.sigma.data();
  }
  if (c.variable) p.diffusion = d.diffusion.data();
  mgmc_params& par = d.par;
  par.nlevel = c.nlevel;
  par.smoother = MGMC_SSOR;
  par.npresmooth = 1;
  par.npostsmooth = 1;
  par.ncoarsesmooth = 1;
  par.omega = 1.0;
  par.cycle = 1;
  par.coarse_scaling = 1.0;
  par.coarse_solver = MGMC_COARSE_CHOLESKY;
  par.precision = c.precision;
  par.nchains = c.K;
  par.seed = 7;
}

void run(const Case& c) {
  std::printf("case %s\n", c.name);
  std::fflush(stdout);
  ProblemData d;
  build(c, d);
  const int64_t N = d.N;
  mgmc_problem& p = d.p;
  mgmc_params& par = d.par;
  mgmc_handle* h = nullptr;
  check(mgmc_create(&p, &par, &h), "mgmc_create");
  if (c.ordering != MGMC_ORDER_MULTICOLOUR) check(mgmc_set_ordering(h, c.ordering), "mgmc_set_ordering");
  // sampling cycles from x = 0 with a shared right-hand side, probes recorded
  std::vector<double> f((size_t)N), x((size_t)(N * c.K), 0.0);
  for (int64_t i = 0; i < N; ++i) f[(size_t)i] = std::sin(0.37 * (double)i);
  check(mgmc_set_rhs(h, f.data(), 0), "mgmc_set_rhs");
  check(mgmc_set_state(h, x.data()), "mgmc_set_state");
  const int64_t probes[3] = {0, N / 2, N - 1};
  check(mgmc_set_probes(h, 3, probes), "mgmc_set_probes");
  const int ncycles = 2;
  std::vector<double> obs((size_t)(ncycles * c.K * 3));
  check(mgmc_run(h, ncycles, obs.data(), 0), "mgmc_run");
  check(mgmc_get_state(h, x.data()), "mgmc_get_state");
  for (double v : x)
    if (!std::isfinite(v)) {
      std::printf("FAIL %s: non-finite sample\n", c.name);
      ++failures;
      break;
    }
  // a multigrid solve of Q x = f (FP64 residuals), which must converge
  std::vector<double> b((size_t)(N * c.K)), xs((size_t)(N * c.K), 0.0), hist(64);
  for (int k = 0; k < c.K; ++k)
    for (int64_t i = 0; i < N; ++i) b[(size_t)(k * N + i)] = f[(size_t)i] * (1.0 + k);
  int niter = -1;
  const int st = mg_solve(h, b.data(), xs.data(), 1e-6, 1e300, 40, hist.data(), &niter);
  if (st != MGMC_OK || !(hist[niter] < 1e-6 * hist[0])) {
    std::printf("FAIL %s: solve status %d after %d iterations, residual ratio %.3e\n", c.name, st, niter,
                niter >= 0 ? hist[niter] / hist[0] : -1.0);
    ++failures;
  } else {
    std::printf("  solve: %d iterations, residual ratio %.2e\n", niter, hist[niter] / hist[0]);
  }
  check(mgmc_destroy(h), "mgmc_destroy");
}

// A batch of 3D and 2D instances of different sizes, precisions and chain
// counts (one with more chains than a low-rank update workgroup serves):
// the merged launches (mgmc_batch_run_dev) must give every instance the
// observables and final state of its own run, bit for bit.
void batch_case() {
  std::printf("case batch of 7 instances, 3d and 2d (merged launches)\n");
  std::fflush(stdout);
  const Case cs[] = {
      {"16^3 fp32", 3, {16, 16, 16}, 3, 3, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"12x16x8 fp64", 3, {12, 16, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 1, false},
      {"24^3 fp32", 3, {24, 24, 24}, 2, 3, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 3, false},
      {"16^3 fp32 no measurements", 3, {16, 16, 16}, 0, 2, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 1, false},
      {"2d 24x20 fp32", 2, {24, 20, 1}, 3, 2, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"2d 32x32 fp64", 2, {32, 32, 1}, 2, 3, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 1, false},
      // more chains than one low-rank update workgroup serves: separate solve launches
      {"8^3 fp32, 9 chains", 3, {8, 8, 8}, 2, 2, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 9, false},
  };
  const int n = 7, ncycles = 2;
  ProblemData d[n];
  mgmc_handle* own[n];
  mgmc_handle* bat[n];
  std::vector<double> obs_own[n], obs_bat[n];
  double* ptrs[n];
  for (int i = 0; i < n; ++i) {
    build(cs[i], d[i]);
    d[i].par.seed = 100 + i;
    check(mgmc_create(&d[i].p, &d[i].par, &own[i]), "mgmc_create");
    check(mgmc_create(&d[i].p, &d[i].par, &bat[i]), "mgmc_create");
    const int64_t N = d[i].N;
    std::vector<double> f((size_t)N), x((size_t)(N * cs[i].K), 0.0);
    for (int64_t v = 0; v < N; ++v) f[(size_t)v] = std::cos(0.29 * (double)v + i);
    const int64_t probes[3] = {1, N / 3, N - 2};
    for (mgmc_handle* h : {own[i], bat[i]}) {
      check(mgmc_set_rhs(h, f.data(), 0), "mgmc_set_rhs");
      check(mgmc_set_state(h, x.data()), "mgmc_set_state");
      check(mgmc_set_probes(h, 3, probes), "mgmc_set_probes");
    }
    obs_own[i].assign((size_t)(ncycles * cs[i].K * 3), 0.0);
    obs_bat[i].assign((size_t)(ncycles * cs[i].K * 3), -1.0);
    // emulated device memory is host memory
    check(mgmc_run_dev(own[i], ncycles, obs_own[i].data(), 0), "mgmc_run_dev");
    ptrs[i] = obs_bat[i].data();
  }
  mgmc_batch* b = nullptr;
  check(mgmc_batch_create(n, bat, &b), "mgmc_batch_create");
  check(mgmc_batch_run_dev(b, ncycles, ptrs, 0), "mgmc_batch_run_dev");
  int64_t merged = 0, separate = 0;
  check(mgmc_batch_stats(b, &merged, &separate), "mgmc_batch_stats");
  std::printf("  %lld launches per cycle merged, %lld separate\n", (long long)merged, (long long)separate);
  if (!(merged < separate)) {
    std::printf("FAIL batch: no launch was merged\n");
    ++failures;
  }
  for (int i = 0; i < n; ++i) {
    if (std::memcmp(obs_own[i].data(), obs_bat[i].data(), obs_own[i].size() * sizeof(double)) != 0) {
      std::printf("FAIL batch: observables of instance %d differ\n", i);
      ++failures;
    }
    const size_t nx = (size_t)(d[i].N * cs[i].K);
    std::vector<double> xo(nx), xb(nx);
    check(mgmc_get_state(own[i], xo.data()), "mgmc_get_state");
    check(mgmc_get_state(bat[i], xb.data()), "mgmc_get_state");
    if (std::memcmp(xo.data(), xb.data(), nx * sizeof(double)) != 0) {
      std::printf("FAIL batch: final state of instance %d differs\n", i);
      ++failures;
    }
  }
  // the persistent form (workgroups looping over the blocks of every
  // instance) continues the same chains: compare with two more own cycles
  check(mgmc_batch_set_persistent(b, 1), "mgmc_batch_set_persistent");
  for (int i = 0; i < n; ++i) {
    check(mgmc_run_dev(own[i], ncycles, obs_own[i].data(), 0), "mgmc_run_dev");
    std::fill(obs_bat[i].begin(), obs_bat[i].end(), -1.0);
  }
  check(mgmc_batch_run_dev(b, ncycles, ptrs, 0), "mgmc_batch_run_dev persistent");
  for (int i = 0; i < n; ++i)
    if (std::memcmp(obs_own[i].data(), obs_bat[i].data(), obs_own[i].size() * sizeof(double)) != 0) {
      std::printf("FAIL batch: persistent observables of instance %d differ\n", i);
      ++failures;
    }
  check(mgmc_batch_destroy(b), "mgmc_batch_destroy");
  for (int i = 0; i < n; ++i) {
    check(mgmc_destroy(own[i]), "mgmc_destroy");
    check(mgmc_destroy(bat[i]), "mgmc_destroy");
  }
  // a handle whose level-1 sweep cannot be merged (the single-pass box march,
  // reached after level 0 has been recorded): the batch is refused and the
  // handle's own cycles are those of an identical handle never offered to a batch
  {
    ProblemData e;
    build(cs[0], e);
    mgmc_handle *hx = nullptr, *hy = nullptr;
    check(mgmc_create(&e.p, &e.par, &hx), "mgmc_create");
    check(mgmc_create(&e.p, &e.par, &hy), "mgmc_create");
    const int64_t N = e.N, probes[2] = {2, N / 2};
    std::vector<double> f((size_t)N), x((size_t)(N * cs[0].K), 0.0);
    for (int64_t v = 0; v < N; ++v) f[(size_t)v] = std::sin(0.13 * (double)v);
    for (mgmc_handle* h : {hx, hy}) {
      check(mgmc_set_tuning(h, "box_march", 1), "mgmc_set_tuning");
      check(mgmc_set_rhs(h, f.data(), 0), "mgmc_set_rhs");
      check(mgmc_set_state(h, x.data()), "mgmc_set_state");
      check(mgmc_set_probes(h, 2, probes), "mgmc_set_probes");
    }
    mgmc_batch* bx = nullptr;
    // W-cycles revisit levels in an order that differs between instances
    {
      ProblemData w;
      build(cs[0], w);
      w.par.cycle = 2;
      mgmc_handle* hw = nullptr;
      check(mgmc_create(&w.p, &w.par, &hw), "mgmc_create");
      if (mgmc_batch_create(1, &hw, &bx) != MGMC_ERR_UNSUPPORTED) {
        std::printf("FAIL batch: a W-cycle handle was accepted\n");
        ++failures;
      }
      check(mgmc_destroy(hw), "mgmc_destroy");
    }
    if (mgmc_batch_create(1, &hx, &bx) != MGMC_ERR_UNSUPPORTED) {
      std::printf("FAIL batch: a handle with an unmergeable kernel was accepted\n");
      ++failures;
    }
    std::vector<double> ox((size_t)(2 * cs[0].K * 2)), oy((size_t)(2 * cs[0].K * 2));
    check(mgmc_run_dev(hx, 2, ox.data(), 0), "mgmc_run_dev");
    check(mgmc_run_dev(hy, 2, oy.data(), 0), "mgmc_run_dev");
    if (std::memcmp(ox.data(), oy.data(), ox.size() * sizeof(double)) != 0) {
      std::printf("FAIL batch: a refused handle's cycles changed\n");
      ++failures;
    }
    check(mgmc_destroy(hx), "mgmc_destroy");
    check(mgmc_destroy(hy), "mgmc_destroy");
  }
}

// The level-0 red-black sweep with its (chain, tile, plane) space split
// evenly over the resident workgroups (rb_balanced: runs that cross from one
// column into the next) must give the samples of its z-segment form, and
// with the coarse correction added while the first post-sweep reads its
// planes, those of a separate prolongation launch (to rounding); the
// residuals' low-rank terms carried from the SMW steps before them, those of
// a second gather over x (to rounding).
void balanced_case() {
  std::printf("case balanced level-0 sweep split\n");
  std::fflush(stdout);
  for (int precision : {MGMC_FP32, MGMC_FP64}) {
    // 8 rows per tile at 64 threads: 2 tiles x 4 chains = 8 columns of 19
    // planes over 6 emulated resident workgroups
    const Case c = {"40x12x20", 3, {40, 12, 20}, 2, 2, precision, MGMC_ORDER_MULTICOLOUR, 4, false};
    ProblemData d;
    build(c, d);
    const int64_t N = d.N;
    std::vector<double> f((size_t)N), x((size_t)(N * c.K), 0.0), st[4], obs[4];
    for (int64_t v = 0; v < N; ++v) f[(size_t)v] = std::sin(0.21 * (double)v);
    const int64_t probes[2] = {3, N / 2};
    for (int v = 0; v < 4; ++v) {
      const int bal = v == 0 ? 0 : 1;
      mgmc_handle* h = nullptr;
      check(mgmc_create(&d.p, &d.par, &h), "mgmc_create");
      check(mgmc_set_tuning(h, "rb_threads", 64), "mgmc_set_tuning");
      check(mgmc_set_tuning(h, "rb_balanced", bal), "mgmc_set_tuning");
      if (v == 2) check(mgmc_set_tuning(h, "fold_prolongation", 0), "mgmc_set_tuning");
      if (v == 3) check(mgmc_set_tuning(h, "lr_carry", 0), "mgmc_set_tuning");
      int got = -1;
      check(mgmc_get_tuning(h, "rb_balanced", &got), "mgmc_get_tuning");
      if (got != bal) {
        std::printf("FAIL balanced: rb_balanced reads %d\n", got);
        ++failures;
      }
      check(mgmc_set_rhs(h, f.data(), 0), "mgmc_set_rhs");
      check(mgmc_set_state(h, x.data()), "mgmc_set_state");
      check(mgmc_set_probes(h, 2, probes), "mgmc_set_probes");
      obs[v].assign((size_t)(3 * c.K * 2), 0.0);
      check(mgmc_run(h, 3, obs[v].data(), 0), "mgmc_run");
      st[v].assign((size_t)(N * c.K), 0.0);
      check(mgmc_get_state(h, st[v].data()), "mgmc_get_state");
      check(mgmc_destroy(h), "mgmc_destroy");
    }
    if (std::memcmp(st[0].data(), st[1].data(), st[0].size() * sizeof(double)) != 0 ||
        std::memcmp(obs[0].data(), obs[1].data(), obs[0].size() * sizeof(double)) != 0) {
      std::printf("FAIL balanced: %s samples differ from the z-segment split\n", precision == MGMC_FP32 ? "fp32" : "fp64");
      ++failures;
    }
    const double tol = precision == MGMC_FP64 ? 1e-12 : 2e-5;
    for (int v = 2; v < 4; ++v) {
      double dmax = 0.0, amax = 0.0;
      for (size_t i = 0; i < st[1].size(); ++i) {
        dmax = std::max(dmax, std::fabs(st[1][i] - st[v][i]));
        amax = std::max(amax, std::fabs(st[v][i]));
      }
      if (!(dmax <= tol * amax)) {
        std::printf("FAIL %s: %s differs from the reference form by %.3e (max |x| %.3e)\n",
                    v == 2 ? "fold" : "carry", precision == MGMC_FP32 ? "fp32" : "fp64", dmax, amax);
        ++failures;
      }
    }
  }
}

}  // namespace

int main() {
  const Case cases[] = {
      {"3d 16^3, 3 measurements, fp64, multicolour", 3, {16, 16, 16}, 3, 3, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"3d 16^3, 3 measurements, fp32, multicolour", 3, {16, 16, 16}, 3, 3, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 2, false},
      // more chains than one low-rank update workgroup serves: separate solve launch
      {"3d 8^3, 2 measurements, fp32, 9 chains", 3, {8, 8, 8}, 2, 2, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 9, false},
      {"3d 12x16x8 (x not a multiple of 8), fp64", 3, {12, 16, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 1, false},
      {"2d 24x20, fp32", 2, {24, 20, 1}, 3, 2, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"3d 8^3, lexicographic, fp64", 3, {8, 8, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_LEXICOGRAPHIC, 1, false},
      {"3d 8^3, variable diffusion, fp64", 3, {8, 8, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 2, true},
  };
  for (const Case& c : cases) run(c);
  batch_case();
  balanced_case();
  if (failures) return 1;
  std::printf("emu_cycle ok\n");
  return 0;
}
