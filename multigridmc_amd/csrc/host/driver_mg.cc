// driver_mg: the reference's multigrid solver driver (src/driver_mg.cc) on the
// GPU library, through the adapter classes (HipMultigridPreconditioner,
// HipLoopSolver): the Richardson loop solver preconditioned by one multigrid
// cycle, with the reference's right-hand side (std::mt19937_64 seeded 1482817,
// standard normal entries), per-iteration residual print-out and solution.vtk.
// gpu.ordering = "lexicographic" sweeps in the reference's order, which
// reproduces its residual table.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <random>

#include "../adapters/mgmc_adapters.hh"
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
  // the operator and the preconditioned loop solver through the classes with
  // the reference's interface shapes (src/driver_mg.cc:150-173)
  mgmc_amd::Lattice lat;
  lat.dim = P.dim;
  for (int d = 0; d < 3; ++d) lat.n[d] = P.n[d];
  std::shared_ptr<mgmc_amd::LinearOperator> op =
      P.fem ? mgmc_amd::LinearOperator::from_csr(lat, P.A_rowptr, P.A_col, P.A_val)
            : mgmc_amd::LinearOperator::shifted_laplace_fd(lat, P.kappa_sq);
  if (P.m > 0) op = mgmc_amd::LinearOperator::with_measurements(op, P.m, P.B_rowptr, P.B_col, P.B_val, P.sigma);
  mgmc_amd::MultigridParameters mp;
  mp.nlevel = multigrid.nlevel;
  mp.smoother = multigrid.smoother;
  mp.npresmooth = multigrid.npresmooth;
  mp.npostsmooth = multigrid.npostsmooth;
  mp.ncoarsesmooth = multigrid.ncoarsesmooth;
  mp.omega = multigrid.omega;
  mp.cycle = multigrid.cycle;
  mp.coarse_scaling = multigrid.coarse_scaling;
  mp.coarse_solver = multigrid.coarse_solver;
  mgmc_amd::DeviceOptions opt;
  opt.precision = gpu.precision == "fp32" ? MGMC_FP32 : MGMC_FP64;
  opt.ordering = gpu.ordering == "lexicographic" ? MGMC_ORDER_LEXICOGRAPHIC : MGMC_ORDER_MULTICOLOUR;
  mgmc_amd::IterativeSolverParameters sp;
  sp.rtol = solver.rtol;
  sp.atol = solver.atol;
  sp.maxiter = solver.maxiter;
  sp.verbose = solver.verbose;
  std::vector<double> b((size_t)P.N), x((size_t)P.N, 0.0);
  try {
    auto pre = std::make_shared<mgmc_amd::HipMultigridPreconditioner>(op, mp, opt);
    mgmc_amd::HipLoopSolver loop(op, pre, sp);
    std::cout << std::endl;
    // right-hand side as in src/driver_mg.cc:163-171
    std::mt19937_64 rng(1482817);
    std::normal_distribution<double> normal_dist(0.0, 1.0);
    for (int64_t ell = 0; ell < P.N; ++ell) b[ell] = normal_dist(rng);
    loop.apply(b, x);
  } catch (const mgmc_amd::Error& e) {
    die(e.what());
  }
  write_vtk("solution.vtk", P, {{"numerical", &x}});
  return 0;
}
