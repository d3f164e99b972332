// Shared host code of the command-line drivers driver_mgmc and driver_mg:
// the reference's configuration groups (src/auxilliary/parameters.cc), the
// problem set-up (lattice, prior, measurements: src/driver_mgmc.cc:91-161),
// the measurement harnesses (src/measurements/*.cc) and the VTK writers
// (src/auxilliary/vtk_writer{2d,3d}.cc), all on top of the C ABI of
// libmgmc_amd.so. Errors print a message and exit(-1), as the reference does.
#ifndef MGMC_HOST_DRIVER_COMMON_H
#define MGMC_HOST_DRIVER_COMMON_H

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/mgmc_amd.h"
#include "config.h"

namespace mgmc_host {

[[noreturn]] void die(const std::string& msg);
void check(int status, const char* what);

struct GeneralParameters {
  int dim = 2;
  bool do_cholesky = false, do_ssor = false, do_multigridmc = false;
  bool measure_time_per_sample = false, save_posterior_statistics = false;
  bool measure_convergence = false, measure_mse = false;
  std::string operator_name = "posterior";
  void parse(const Setting& root);
};

struct LatticeParameters {
  int nx = 0, ny = 0, nz = 0;
  void parse(const Setting& root);
};

struct CholeskyParameters {
  std::string factorisation = "sparse";
  void parse(const Setting& root);
};

struct SmootherParameters {
  double omega = 1.0;
  int nsmooth = 1;
  void parse(const Setting& root);
};

struct IterativeSolverParameters {
  double rtol = 1e-12, atol = 1e-15;
  int maxiter = 100, verbose = 2;
  void parse(const Setting& root);
};

struct MultigridParameters {
  int nlevel = 4;
  std::string smoother = "SSOR", coarse_solver = "Cholesky";
  int npresmooth = 1, npostsmooth = 1, ncoarsesmooth = 1, cycle = 1, verbose = 0;
  double omega = 1.0, coarse_scaling = 1.0;
  void parse(const Setting& root);
};

struct SamplingParameters {
  int nsamples = 0, nwarmup = 0;
  int nstepsconvergence = 0, nsamplesconvergence = 0, nthreadsconvergence = 1;
  int nstepsmse = 0, nsamplesmse = 0;
  void parse(const Setting& root);
};

struct PriorParameters {
  std::string pde_model = "shiftedlaplace_fd", correlationlength_model = "constant";
  double Lambda = 1.0, Lambda_min = 1.0, Lambda_max = 1.0;
  void parse(const Setting& root);
};

struct MeasurementParameters {
  double radius = 0.0, variance_scaling = 1.0, variance_global = 0.0, mean_global = 0.0;
  bool measure_global = false;
  std::string filename;
  std::vector<double> sample_location;
  int dim = 0, n = 0;
  std::vector<double> locations;  // n * dim
  std::vector<double> mean, variance;
  void parse(const Setting& root);
};

// Optional group "gpu" (no counterpart in the reference's files): chains per
// handle, precision of the sampling cycles, Philox seed.
struct GpuParameters {
  int nchains = 1;               // chains of the timed sampling run
  int convergence_chains = 256;  // chains per batch of the convergence / MSE runs
  std::string precision = "fp64";
  std::string ordering = "multicolour";  // or "lexicographic": the reference's sweep order
  uint64_t seed = 1;
  void parse(const Setting& root);
};

// Posterior (or prior) problem on the lattice in the reference's ordering.
struct Problem {
  int dim = 2;
  int n[3] = {1, 1, 1};
  int64_t N = 0;
  double kappa_sq = 0.0;
  std::vector<double> kappa_field;  // per vertex (periodic correlation-length model), empty: constant
  bool fem = false;
  std::vector<int64_t> A_rowptr, A_col;  // FEM prior as CSR
  std::vector<double> A_val;
  int m = 0;
  std::vector<int64_t> B_rowptr, B_col;
  std::vector<double> B_val, sigma, y;
  mgmc_problem view() const;
};

Problem build_problem(const GeneralParameters& gp, const LatticeParameters& lp, const PriorParameters& pp,
                      const MeasurementParameters& mp);

// Sparse observation vector b (measurement_vector at the sample location).
struct Observation {
  std::vector<int64_t> idx;
  std::vector<double> val;
};
Observation measurement_vector(const Problem& P, const std::vector<double>& x0, double radius);

struct SamplerSpec;

// Handle owning wrapper.
struct Handle {
  mgmc_handle* h = nullptr;
  int nchains = 1;
  Handle(const Problem& P, const mgmc_params& par);
  Handle(const SamplerSpec& s, int nchains);
  ~Handle();
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
};

// Exact statistics: x_post = Q^{-1} B Sigma^{-1} y (prior mean zero) and
// mean / variance of z = b^T x from FP64 multigrid solves
// (LinearOperator::mean, observed_mean_and_variance, linear_operator.hh:131-189).
struct Exact {
  std::vector<double> x_post, f;  // f = Q x_post, the sampler's right-hand side
  double mean = 0.0, variance = 0.0;
};
Exact exact_statistics(const Problem& P, const MultigridParameters& mg, const Observation& b);

// A sampler: the problem and the parameters of its handles (the harnesses
// create handles with the chain counts they need, as the reference deep-copies
// its sampler per thread).
struct SamplerSpec {
  const Problem* P;
  mgmc_params par;
  int cholesky = -1;  // mgmc_factorisation of the host Cholesky sampler; -1: a GPU handle
  std::shared_ptr<mgmc_cholesky> factor;  // its factorisation, shared by every set of chains
};
SamplerSpec multigridmc_sampler(const Problem& P, const MultigridParameters& mg, const GpuParameters& gpu);
SamplerSpec ssor_sampler(const Problem& P, const SmootherParameters& sm, const GpuParameters& gpu);
// the reference's sparse / lowrank / dense Cholesky sampler (host code of the
// library); factorises once and prints the time as src/driver_mgmc.cc:174-192
SamplerSpec cholesky_sampler(const Problem& P, const CholeskyParameters& ch, const GpuParameters& gpu);

// K chains of one sampler behind the harnesses: a GPU handle (MGMC, SSOR) or
// the host Cholesky sampler, whose samples are independent.
struct Chains {
  int nchains = 1;
  virtual ~Chains() = default;
  // right-hand side Q x_post, zero states, observable b
  virtual void prepare(const Exact& ex, const Observation& b) = 0;
  // nsteps samples per chain; obs (nsteps x nchains, may be NULL) receives b.x
  virtual void run(int nsteps, double* obs) = 0;
  // pointwise mean and variance over nsamples steps of all chains
  virtual void moments(int nsamples, double* mean, double* var) = 0;
};
std::unique_ptr<Chains> make_chains(const SamplerSpec& s, int nchains);

// Harnesses (src/measurements/*.cc).
void measure_sampling_time(const SamplerSpec& s, const Exact& ex, const Observation& b, const SamplingParameters& sp,
                           const GpuParameters& gpu, const std::string& label, const std::string& filename);
void measure_convergence(const SamplerSpec& s, const Exact& ex, const Observation& b, const SamplingParameters& sp,
                         const GpuParameters& gpu, const std::string& filename);
void measure_mean_squared_error(const SamplerSpec& s, const Exact& ex, const Observation& b,
                                const SamplingParameters& sp, const GpuParameters& gpu, const std::string& filename);
void posterior_statistics(const SamplerSpec& s, const Exact& ex, const SamplingParameters& sp,
                          const MeasurementParameters& mp, const GpuParameters& gpu);

// VTK structured-points writer (vtk_writer2d.cc / vtk_writer3d.cc).
void write_vtk(const std::string& filename, const Problem& P,
               const std::vector<std::pair<std::string, const std::vector<double>*>>& states);
void write_vtk_circle(const std::vector<double>& centre, double radius, const std::string& filename);

}  // namespace mgmc_host

#endif
