// Runs libmgmc_amd's kernels on the CPU (hip_emu.h) under the sanitizers:
// sampling cycles (fused red-black sweeps with the folded coarse correction,
// box-level plane sweeps, low-rank kernels, the coarsest-level sample) and
// multigrid solves, in FP64 and FP32, multicolour and lexicographic order,
// with constant and per-vertex stencils, on lattices whose x extent is and is
// not a multiple of 8. A solve must converge; the samples must be finite.
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

void run(const Case& c) {
  std::printf("case %s\n", c.name);
  std::fflush(stdout);
  int64_t N = 1;
  for (int d = 0; d < c.dim; ++d) N *= c.n[d] - 1;
  // measurements: balls of radius 0.2 at seeded centres
  std::vector<std::vector<std::pair<int64_t, double>>> rows((size_t)N);
  std::vector<double> sigma;
  for (int k = 0; k < c.m; ++k) {
    double x0[3] = {0.3 + 0.17 * k, 0.6 - 0.11 * k, 0.45};
    std::vector<int64_t> idx((size_t)N);
    std::vector<double> val((size_t)N);
    int64_t nnz = 0;
    check(mgmc_measurement_vector(c.dim, c.n, x0, 0.2, N, idx.data(), val.data(), &nnz), "measurement_vector");
    for (int64_t e = 0; e < nnz; ++e) rows[(size_t)idx[e]].push_back({k, val[e]});
    sigma.push_back(1e-3 * (1.0 + k));
  }
  std::vector<int64_t> rp(1, 0), col;
  std::vector<double> bval;
  for (auto& r : rows) {
    for (auto& e : r) {
      col.push_back(e.first);
      bval.push_back(e.second);
    }
    rp.push_back((int64_t)col.size());
  }
  int64_t ncell = 1;
  for (int d = 0; d < c.dim; ++d) ncell *= c.n[d];
  std::vector<double> diffusion((size_t)ncell);
  for (int64_t i = 0; i < ncell; ++i) diffusion[(size_t)i] = 1.0 + 9.0 * (double)((i * 7919) % 13) / 12.0;
  mgmc_problem p{};
  p.dim = c.dim;
  for (int d = 0; d < 3; ++d) p.n[d] = c.n[d];
  p.kappa_sq = 10.0;
  p.m = c.m;
  if (c.m > 0) {
    p.B_rowptr = rp.data();
    p.B_col = col.data();
    p.B_val = bval.data();
    p.sigma = sigma.data();
  }
  if (c.variable) p.diffusion = diffusion.data();
  mgmc_params par{};
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

}  // namespace

int main() {
  const Case cases[] = {
      {"3d 16^3, 3 measurements, fp64, multicolour", 3, {16, 16, 16}, 3, 3, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"3d 16^3, 3 measurements, fp32, multicolour", 3, {16, 16, 16}, 3, 3, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"3d 12x16x8 (x not a multiple of 8), fp64", 3, {12, 16, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 1, false},
      {"2d 24x20, fp32", 2, {24, 20, 1}, 3, 2, MGMC_FP32, MGMC_ORDER_MULTICOLOUR, 2, false},
      {"3d 8^3, lexicographic, fp64", 3, {8, 8, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_LEXICOGRAPHIC, 1, false},
      {"3d 8^3, variable diffusion, fp64", 3, {8, 8, 8}, 2, 2, MGMC_FP64, MGMC_ORDER_MULTICOLOUR, 2, true},
  };
  for (const Case& c : cases) run(c);
  if (failures) return 1;
  std::printf("emu_cycle ok\n");
  return 0;
}
