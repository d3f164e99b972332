// driver_mgmc: the reference's sampling driver (src/driver_mgmc.cc) on the
// GPU library. Reads the same configuration and measurement files, runs the
// Cholesky / SSOR / multigrid MC samplers through the C ABI and writes the same
// files (timeseries_*.txt, convergence_*.txt, meansquarederror_*.txt,
// posterior.vtk). An optional "gpu" group sets chains, precision and seed.
#include <chrono>
#include <cstdio>
#include <ctime>
#include <iostream>

#include "driver_common.h"

using namespace mgmc_host;

int main(int argc, char* argv[]) {
  if (argc != 2) {
    std::cout << "Usage: " << argv[0] << " CONFIGURATIONFILE" << std::endl;
    return -1;
  }
  auto t_start = std::chrono::system_clock::now();
  std::time_t start_time = std::chrono::system_clock::to_time_t(t_start);
  std::cout << std::endl;
  std::cout << "+--------------------------------+" << std::endl;
  std::cout << "! Multigrid Monte Carlo sampling !" << std::endl;
  std::cout << "+--------------------------------+" << std::endl;
  std::cout << std::endl;
  std::cout << "Starting run at " << std::ctime(&start_time) << std::endl;
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
  CholeskyParameters cholesky;
  SmootherParameters smoother;
  MultigridParameters multigrid;
  SamplingParameters sampling;
  PriorParameters prior;
  MeasurementParameters measurements;
  GpuParameters gpu;
  general.parse(root);
  lattice.parse(root);
  cholesky.parse(root);
  smoother.parse(root);
  multigrid.parse(root);
  sampling.parse(root);
  prior.parse(root);
  measurements.parse(root);
  gpu.parse(root);
  if (measurements.dim != general.dim) die("dimension of measurement locations differs from problem dimension");

  Problem P = build_problem(general, lattice, prior, measurements);
  Observation b = measurement_vector(P, measurements.sample_location, measurements.radius);
  Exact ex = exact_statistics(P, multigrid, b);

  struct Run {
    bool enabled;
    const char* title;
    const char* label;
    const char* suffix;
    SamplerSpec spec;
  };
  Run runs[3] = {
      {general.do_cholesky, "**** Cholesky ****", "Cholesky", "cholesky",
       general.do_cholesky ? cholesky_sampler(P, cholesky, gpu) : SamplerSpec{&P, {}}},
      {general.do_ssor, "**** SSOR ****", "SSOR", "ssor", ssor_sampler(P, smoother, gpu)},
      {general.do_multigridmc, "**** Multigrid MC ****", "MGMC", "multigridmc", multigridmc_sampler(P, multigrid, gpu)},
  };
  std::cout << std::endl;
  for (int r = 0; r < 3; ++r) {
    if (!runs[r].enabled) continue;
    std::cout << runs[r].title << std::endl;
    const std::string sfx = runs[r].suffix;
    if (general.measure_time_per_sample) {
      measure_sampling_time(runs[r].spec, ex, b, sampling, gpu, runs[r].label, "timeseries_" + sfx + ".txt");
      std::cout << std::endl;
    }
    if (general.measure_convergence) {
      measure_convergence(runs[r].spec, ex, b, sampling, gpu, "convergence_" + sfx + ".txt");
      std::cout << std::endl;
    }
    if (r == 2 && general.save_posterior_statistics) {
      posterior_statistics(runs[r].spec, ex, sampling, measurements, gpu);
      std::cout << std::endl;
    }
    if (general.measure_mse) {
      measure_mean_squared_error(runs[r].spec, ex, b, sampling, gpu, "meansquarederror_" + sfx + ".txt");
      std::cout << std::endl;
    }
  }
  auto t_finish = std::chrono::system_clock::now();
  std::time_t finish_time = std::chrono::system_clock::to_time_t(t_finish);
  unsigned int elapsed = (unsigned int)std::chrono::duration<double>(t_finish - t_start).count();
  int seconds = elapsed % 60;
  elapsed /= 60;
  int minutes = elapsed % 60;
  elapsed /= 60;
  int hours = elapsed;
  std::cout << "Completed run at " << std::ctime(&finish_time);
  printf("Total elapsed time = %3d h %2d m %2d s\n", hours, minutes, seconds);
  std::cout << std::endl;
  return 0;
}
