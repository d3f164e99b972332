// driver_mg: the reference's multigrid solver driver (src/driver_mg.cc) on the
// GPU library: the Richardson loop solver preconditioned by one multigrid
// cycle, with the reference's right-hand side (std::mt19937_64 seeded 1482817,
// standard normal entries), per-iteration residual print-out and solution.vtk.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <random>

#include "driver_common.h"

using namespace mgmc_host;

int main(int argc, char* argv[]) {
  if (argc != 2) {
    std::cout << "Usage: " << argv[0] << " CONFIGURATIONFILE" << std::endl;
    return -1;
  }
  std::cout << std::endl;
  std::cout << "+------------------+" << std::endl;
  std::cout << "! Multigrid solver !" << std::endl;
  std::cout << "+------------------+" << std::endl;
  std::cout << std::endl;
  std::string filename(argv[1]);
  std::cout << "Reading parameters from file '" << filename << "'" << std::endl;
  Setting root;
  try {
    root = read_config_file(filename);
  } catch (const ConfigError& e) {
    die(e.what());
  }
  GeneralParameters general;
  LatticeParameters lattice;
  SmootherParameters smoother;
  MultigridParameters multigrid;
  IterativeSolverParameters solver;
  PriorParameters prior;
  MeasurementParameters measurements;
  GpuParameters gpu;
  general.parse(root);
  lattice.parse(root);
  smoother.parse(root);
  multigrid.parse(root);
  solver.parse(root);
  prior.parse(root);
  measurements.parse(root);
  gpu.parse(root);
  if (measurements.dim != general.dim) die("dimension of measurement locations differs from problem dimension");
  Problem P = build_problem(general, lattice, prior, measurements);
  SamplerSpec spec = multigridmc_sampler(P, multigrid, gpu);
  // the reference's preconditioner always solves the coarsest level with
  // Cholesky (multigrid_preconditioner.cc:41-45)
  if (multigrid.coarse_solver != "Cholesky")
    std::cout << "WARNING: ignoring coarse solver setting '" << multigrid.coarse_solver << "', using Cholesky."
              << std::endl;
  spec.par.coarse_solver = MGMC_COARSE_CHOLESKY;
  Handle h(spec, 1);
  std::cout << std::endl;
  // right-hand side as in src/driver_mg.cc:163-171
  std::vector<double> b((size_t)P.N), x((size_t)P.N, 0.0);
  std::mt19937_64 rng(1482817);
  std::normal_distribution<double> normal_dist(0.0, 1.0);
  for (int64_t ell = 0; ell < P.N; ++ell) b[ell] = normal_dist(rng);
  std::vector<double> hist((size_t)std::max(solver.maxiter, 1));
  int niter = -1;
  int st = mg_solve(h.h, b.data(), x.data(), solver.rtol, solver.atol, solver.maxiter, hist.data(), &niter);
  if (st != MGMC_OK && st != MGMC_ERR_NOCONV) check(st, "mg_solve");
  // LoopSolver::apply print-out (src/solver/loop_solver.cc:12-53)
  double r0 = 0.0;
  for (double v : b) r0 += v * v;
  r0 = std::sqrt(r0);
  const int nrec = niter >= 0 ? niter + 1 : solver.maxiter;
  if (solver.verbose >= 2) {
    printf("Initial residual ||r_0|| =  %12.4f\n", r0);
    printf("%5s   %8s   %12s   %6s\n", "iter", "||r||", "||r||/||r_0||", "rho");
    double rold = r0;
    for (int k = 0; k < nrec; ++k) {
      printf("%5d   %8.3e   %12.3e   %6.3f\n", k, hist[k], hist[k] / r0, hist[k] / rold);
      rold = hist[k];
    }
  }
  if (solver.verbose >= 1) {
    if (niter >= 0)
      printf("Solver converged after %5d iterations\n||r|| = %8.3e, ||r||/||r_0|| = %8.3e\n", niter, hist[niter],
             hist[niter] / r0);
    else
      printf("Solver failed to converge after %5d iterations\n", solver.maxiter);
  }
  write_vtk("solution.vtk", P, {{"numerical", &x}});
  return 0;
}
