// C++ classes with the reference's interface shapes over the C ABI of
// libmgmc_amd.so (include/mgmc_amd.h), without Eigen:
//
//   Sampler                  src/sampler/sampler.hh:23-78
//   HipMultigridMCSampler    MultigridMCSampler, src/sampler/multigridmc_sampler.hh
//   Preconditioner           src/preconditioner/preconditioner.hh:17-37
//   HipMultigridPreconditioner  MultigridPreconditioner, src/preconditioner/multigrid_preconditioner.hh
//   HipLoopSolver            LoopSolver, src/solver/loop_solver.hh
//   Smoother                 src/smoother/smoother.hh:18-37
//   HipSORSmoother, HipSSORSmoother  SORSmoother, SSORSmoother
//                            (src/smoother/sor_smoother.hh, ssor_smoother.hh)
//   HipSORSampler, HipSSORSampler    SORSampler, SSORSampler
//                            (src/sampler/sor_sampler.hh, ssor_sampler.hh)
//   HipIntergridOperator     IntergridOperatorLinear's restrict / prolongate_add
//                            (src/intergrid/intergrid_operator.hh:74-120)
//   MultigridParameters      src/auxilliary/parameters.hh:150-179
//   IterativeSolverParameters src/auxilliary/parameters.hh:110-133
//   RandomGenerator          src/auxilliary/parallel_random.hh:17-55 (draw_int only)
//   LinearOperator           the data of src/linear_operator/linear_operator.hh:28-213
//                            (lattice, A, B, Sigma) with the ShiftedLaplaceFDOperator
//                            and MeasuredOperator constructors
//
// Vectors are passed as views (ConstVec, Vec) that bind to anything with
// data() and size(): Eigen::VectorXd in the reference, std::vector<double>
// elsewhere. As in the reference, x of Sampler::apply has N values (one
// chain). Errors throw mgmc_amd::Error (status code and message of the
// library) where the reference prints and calls exit(-1).
#ifndef MGMC_ADAPTERS_HH
#define MGMC_ADAPTERS_HH

#include <cstddef>
#include <cstdint>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mgmc_amd.h"

namespace mgmc_amd {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Read-only / writable views of a contiguous FP64 vector.
struct ConstVec {
  const double* p;
  size_t n;
  ConstVec(const double* p_, size_t n_) : p(p_), n(n_) {}
  template <class V>
  ConstVec(const V& v) : p(v.data()), n((size_t)v.size()) {}
};
struct Vec {
  double* p;
  size_t n;
  Vec(double* p_, size_t n_) : p(p_), n(n_) {}
  template <class V>
  Vec(V& v) : p(v.data()), n((size_t)v.size()) {}
};

// Source of a sampler's noise key: the reference's RandomGenerator::draw_int.
// A sampler draws two integers from it at construction (the 64-bit Philox
// key of its chain), so deep copies with different generators get
// independent noise, as the reference's thread-local generators do.
class RandomGenerator {
 public:
  virtual ~RandomGenerator() = default;
  virtual int32_t draw_int() = 0;
};
class Mt19937Generator : public RandomGenerator {
 public:
  explicit Mt19937Generator(uint32_t seed) : g(seed) {}
  int32_t draw_int() override { return (int32_t)g(); }

 private:
  std::mt19937 g;
};

struct MultigridParameters {
  int nlevel = 4;
  std::string smoother = "SSOR";        // "SOR" or "SSOR"
  int npresmooth = 1, npostsmooth = 1, ncoarsesmooth = 1;
  double omega = 1.0;
  int cycle = 1;                        // 1 = V, 2 = W
  double coarse_scaling = 1.0;
  std::string coarse_solver = "Cholesky";  // "Cholesky" or "SSOR"
  int verbose = 0;
};

struct IterativeSolverParameters {
  double rtol = 1e-12, atol = 1e-15;
  int maxiter = 100, verbose = 2;
};

// Options of the GPU path that the reference does not have.
struct DeviceOptions {
  int precision = MGMC_FP64;              // cycles in FP64 or FP32
  int ordering = MGMC_ORDER_MULTICOLOUR;  // or MGMC_ORDER_LEXICOGRAPHIC
  void* stream = nullptr;                 // hipStream_t, NULL: one per handle
};

// Lattice with n[d] cells per direction; unknowns on the interior vertices.
struct Lattice {
  int dim = 2;
  int n[3] = {1, 1, 1};
  int64_t Nvertex() const;
};

// The data of the reference's LinearOperator: Q = A + B Sigma^{-1} B^T.
class LinearOperator {
 public:
  // ShiftedLaplaceFDOperator with constant kappa^2 = 1 / Lambda^2
  // (src/linear_operator/shiftedlaplace_fd_operator.cc:9-57).
  static std::shared_ptr<LinearOperator> shifted_laplace_fd(const Lattice& lattice, double kappa_sq);
  // A prior given as a CSR matrix in reference vertex order (e.g. the FEM
  // operator); the library requires a constant-coefficient stencil.
  static std::shared_ptr<LinearOperator> from_csr(const Lattice& lattice, std::vector<int64_t> rowptr,
                                                  std::vector<int64_t> col, std::vector<double> val);
  // MeasuredOperator (src/linear_operator/measured_operator.cc:9-49): the
  // prior plus ball-average measurements at `centres` (m x dim), Sigma =
  // variance_scaling * variance.
  // profile 0: the reference's top-hat weight; 1: Gaussian weight of the
  // given width (mgmc_measurement_vector_weighted).
  static std::shared_ptr<LinearOperator> measured(const std::shared_ptr<LinearOperator>& prior,
                                                  const std::vector<double>& centres, double radius,
                                                  const std::vector<double>& variance, double variance_scaling = 1.0,
                                                  int profile = 0, double width = 0.5);
  // The prior plus m measurements given as B (N x m CSR, reference vertex
  // order) and Sigma (variances already scaled).
  static std::shared_ptr<LinearOperator> with_measurements(const std::shared_ptr<LinearOperator>& prior, int m,
                                                           std::vector<int64_t> B_rowptr, std::vector<int64_t> B_col,
                                                           std::vector<double> B_val, std::vector<double> sigma);
  const Lattice& get_lattice() const { return lattice; }
  int get_m_lowrank() const { return m; }
  // B^T x (m values) and B Sigma^{-1} y (N values) on the host
  std::vector<double> BT(ConstVec x) const;
  std::vector<double> B_Sigma_inv(ConstVec y) const;
  mgmc_problem problem() const;

 private:
  Lattice lattice;
  double kappa_sq = 0.0;
  std::vector<int64_t> A_rowptr, A_col;
  std::vector<double> A_val;
  int m = 0;
  std::vector<int64_t> B_rowptr, B_col;  // N x m CSR
  std::vector<double> B_val, sigma;
};

// Owns one library handle.
class Handle {
 public:
  Handle(const LinearOperator& op, const MultigridParameters& p, int coarse_solver, const DeviceOptions& o,
         uint64_t seed);
  ~Handle();
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  mgmc_handle* get() const { return h; }

 private:
  mgmc_handle* h = nullptr;
};

class Sampler {
 public:
  Sampler(std::shared_ptr<LinearOperator> linear_operator_, std::shared_ptr<RandomGenerator> rng_)
      : linear_operator(std::move(linear_operator_)), rng(std::move(rng_)) {}
  virtual ~Sampler() = default;
  virtual std::shared_ptr<Sampler> deep_copy(std::shared_ptr<RandomGenerator> rng) = 0;
  // x: the chain's current state on entry, the new sample on exit (N values)
  virtual void apply(ConstVec f, Vec x) const = 0;
  std::shared_ptr<LinearOperator> get_linear_operator() const { return linear_operator; }
  virtual void fix_rhs(ConstVec f) { (void)f; }
  virtual void unfix_rhs() {}

 protected:
  const std::shared_ptr<LinearOperator> linear_operator;
  std::shared_ptr<RandomGenerator> rng;
};

// One MGMC chain on the GPU (MultigridMCSampler::apply,
// src/sampler/multigridmc_sampler.cc:139-145).
class HipMultigridMCSampler : public Sampler {
 public:
  HipMultigridMCSampler(std::shared_ptr<LinearOperator> linear_operator, std::shared_ptr<RandomGenerator> rng,
                        const MultigridParameters& params, const DeviceOptions& options = DeviceOptions());
  // a new handle with the same operator and parameters and a key drawn from `rng`
  std::shared_ptr<Sampler> deep_copy(std::shared_ptr<RandomGenerator> rng) override;
  void apply(ConstVec f, Vec x) const override;
  // f stays on the device; apply() then uses it and ignores its f argument,
  // as the reference's samplers with a fixed right-hand side do
  void fix_rhs(ConstVec f) override;
  void unfix_rhs() override { rhs_fixed = false; }
  uint64_t seed() const { return key; }
  mgmc_handle* handle() const { return h->get(); }

 private:
  MultigridParameters params;
  DeviceOptions options;
  uint64_t key;
  std::unique_ptr<Handle> h;
  bool rhs_fixed = false;
};

class Preconditioner {
 public:
  explicit Preconditioner(std::shared_ptr<LinearOperator> linear_operator_)
      : linear_operator(std::move(linear_operator_)) {}
  virtual ~Preconditioner() = default;
  virtual void apply(ConstVec b, Vec x) = 0;

 protected:
  std::shared_ptr<LinearOperator> linear_operator;
};

// One multigrid cycle from a zero initial guess
// (MultigridPreconditioner::apply, src/preconditioner/multigrid_preconditioner.cc:104-109).
// The coarsest level is always solved with Cholesky (:41-45).
class HipMultigridPreconditioner : public Preconditioner {
 public:
  HipMultigridPreconditioner(std::shared_ptr<LinearOperator> linear_operator, const MultigridParameters& params,
                             const DeviceOptions& options = DeviceOptions());
  void apply(ConstVec b, Vec x) override;
  mgmc_handle* handle() const { return h->get(); }

 private:
  std::unique_ptr<Handle> h;
};

// LoopSolver::apply (src/solver/loop_solver.cc:9-54): x = 0, then
// x -= P(A x - b) until ||r|| / ||r_0|| < rtol and ||r|| < atol, residuals
// and updates in FP64 on the GPU, with the reference's print-out at
// verbose >= 2 (per iteration) and >= 1 (final line).
class HipLoopSolver {
 public:
  HipLoopSolver(std::shared_ptr<LinearOperator> linear_operator,
                std::shared_ptr<HipMultigridPreconditioner> preconditioner, const IterativeSolverParameters& params);
  void apply(ConstVec b, Vec x);
  int niter() const { return niter_; }              // converged iteration, -1 if not converged
  const std::vector<double>& history() const { return hist; }  // ||r_k|| for k = 0..

 private:
  std::shared_ptr<LinearOperator> linear_operator;
  std::shared_ptr<HipMultigridPreconditioner> preconditioner;
  IterativeSolverParameters params;
  int niter_ = -1;
  std::vector<double> hist;
};

// Sweep direction of the SOR smoothers and samplers (src/smoother/sor_smoother.hh).
enum class Direction { forward, backward };

class Smoother {
 public:
  explicit Smoother(std::shared_ptr<LinearOperator> linear_operator_) : linear_operator(std::move(linear_operator_)) {}
  virtual ~Smoother() = default;
  // x: the initial guess on entry, the smoothed solution of Q x = b on exit
  virtual void apply(ConstVec b, Vec x) const = 0;

 protected:
  std::shared_ptr<LinearOperator> linear_operator;
};

// SORSmoother::apply (src/smoother/sor_smoother.cc:41-77) on the GPU: nsmooth
// times (nsmooth sparse sweeps in `direction`, then the SMW correction for
// the measurements), in the device ordering of DeviceOptions (multicolour
// unless lexicographic is asked for).
class HipSORSmoother : public Smoother {
 public:
  HipSORSmoother(std::shared_ptr<LinearOperator> linear_operator, double omega, int nsmooth, Direction direction,
                 const DeviceOptions& options = DeviceOptions());
  void apply(ConstVec b, Vec x) const override;

 private:
  int nsmooth;
  Direction direction;
  std::unique_ptr<Handle> h;
};

// SSORSmoother::apply (src/smoother/ssor_smoother.cc:9-15): nsmooth times a
// forward and a backward SOR sweep with their SMW corrections.
class HipSSORSmoother : public Smoother {
 public:
  HipSSORSmoother(std::shared_ptr<LinearOperator> linear_operator, double omega, int nsmooth,
                  const DeviceOptions& options = DeviceOptions());
  void apply(ConstVec b, Vec x) const override;

 private:
  int nsmooth;
  std::unique_ptr<Handle> h;
};

// SORSampler / SSORSampler::apply (src/sampler/sor_sampler.cc:37-59,
// ssor_sampler.cc:9-15): Gibbs sweeps of N(Q^{-1} f, Q^{-1}) with fresh
// Philox noise on every call, keyed by two draws of `rng`.
class HipSORSampler : public Sampler {
 public:
  HipSORSampler(std::shared_ptr<LinearOperator> linear_operator, std::shared_ptr<RandomGenerator> rng, double omega,
                int nsmooth, Direction direction, const DeviceOptions& options = DeviceOptions());
  std::shared_ptr<Sampler> deep_copy(std::shared_ptr<RandomGenerator> rng) override;
  void apply(ConstVec f, Vec x) const override;
  uint64_t seed() const { return key; }
  mgmc_handle* handle() const { return h->get(); }

 protected:
  HipSORSampler(std::shared_ptr<LinearOperator> linear_operator, std::shared_ptr<RandomGenerator> rng, double omega,
                int nsmooth, Direction direction, const DeviceOptions& options, bool symmetric);
  double omega;
  int nsmooth;
  Direction direction;
  DeviceOptions options;
  bool symmetric;
  uint64_t key;
  std::unique_ptr<Handle> h;
};
class HipSSORSampler : public HipSORSampler {
 public:
  HipSSORSampler(std::shared_ptr<LinearOperator> linear_operator, std::shared_ptr<RandomGenerator> rng, double omega,
                 int nsmooth, const DeviceOptions& options = DeviceOptions());
  std::shared_ptr<Sampler> deep_copy(std::shared_ptr<RandomGenerator> rng) override;
};

// The intergrid operator of a fine lattice and its coarsening
// (IntergridOperatorLinear, src/intergrid/intergrid_operator_linear.cc:8-31):
// restrict(x, x_coarse) sets x_coarse = P^T x, prolongate_add(alpha,
// x_coarse, x) adds alpha P x_coarse to x (intergrid_operator.hh:74-120).
class HipIntergridOperator {
 public:
  explicit HipIntergridOperator(std::shared_ptr<LinearOperator> fine_operator,
                                const DeviceOptions& options = DeviceOptions());
  void restrict(ConstVec x, Vec x_coarse) const;
  void prolongate_add(double alpha, ConstVec x_coarse, Vec x) const;
  int64_t coarse_ndof() const { return nc; }

 private:
  int64_t n = 0, nc = 0;
  std::unique_ptr<Handle> h;
};

}  // namespace mgmc_amd

#endif
