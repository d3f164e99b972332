/* mgmc_amd.h - C ABI of the MI355X (gfx950) multigrid Monte Carlo library.
 *
 * The library implements the MultigridMC hot path behind plain C entry
 * points: the multigrid Monte Carlo (MGMC) sampler and its zero-noise twin,
 * the multigrid preconditioner with the Richardson loop solver. Host code
 * above this ABI mirrors the reference's class interfaces: the C++ classes
 * of multigridmc_amd/csrc/adapters/mgmc_adapters.hh (Sampler,
 * HipMultigridMCSampler, Preconditioner, HipMultigridPreconditioner,
 * HipLoopSolver; libmgmc_adapters.so) and the Python package multigridmc_amd.
 *
 * Conventions
 *  - Every vector handed across the ABI is FP64 in the reference's vertex
 *    order: interior vertices only, x fastest (src/lattice/lattice3d.hh:132-141).
 *    Batched arguments are [chain][vertex] rows.
 *  - Pointers are host pointers unless the name ends in _dev.
 *  - Every function returns 0 on success and a nonzero mgmc_status otherwise;
 *    mgmc_last_error() gives a message. Nothing in the library calls exit().
 *  - One handle per host thread; work is issued on the handle's HIP stream and
 *    host-pointer functions synchronise that stream before returning.
 */
#ifndef MGMC_AMD_H
#define MGMC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgmc_handle mgmc_handle;

enum mgmc_status {
  MGMC_OK = 0,
  MGMC_ERR_ARG = 1,         /* invalid argument (sizes, levels, parameters)   */
  MGMC_ERR_UNSUPPORTED = 2, /* operator the kernels do not cover              */
  MGMC_ERR_NOT_SPD = 3,     /* coarse matrix not positive definite            */
  MGMC_ERR_HIP = 4,         /* HIP runtime error                              */
  MGMC_ERR_NOCONV = 5,      /* solver reached maxiter                         */
  MGMC_ERR_STATE = 6        /* call out of order (e.g. no probes set)         */
};

enum mgmc_smoother { MGMC_SOR = 0, MGMC_SSOR = 1 };
enum mgmc_coarse { MGMC_COARSE_CHOLESKY = 0, MGMC_COARSE_SSOR = 1 };
enum mgmc_precision { MGMC_FP64 = 0, MGMC_FP32 = 1 };
enum mgmc_ordering { MGMC_ORDER_MULTICOLOUR = 0, MGMC_ORDER_LEXICOGRAPHIC = 1 };

/* Posterior precision Q = A + B Sigma^{-1} B^T on the fine lattice
 * (src/linear_operator/linear_operator.hh:28-213).
 * A: if A_rowptr is NULL the library assembles the finite-difference operator
 * -div(a grad) + kappa^2 scaled by the cell volume
 * (ShiftedLaplaceFDOperator, src/linear_operator/shiftedlaplace_fd_operator.cc:9-57):
 * kappa^2 is kappa_sq, or per vertex kappa_sq_field (reference order; the
 * reference's correlation-length models, src/linear_operator/correlationlength_model.hh);
 * a is 1, or per cell `diffusion` (prod n_d values, cell (i, j, k) at
 * (k n_y + j) n_x + i), the coupling across an edge being the mean of a over
 * the cells sharing it. Otherwise A is taken as given (CSR, symmetric,
 * nearest-neighbour couplings, Dirichlet elimination); a constant stencil
 * runs on the fast kernels, varying rows on per-vertex stencils.
 * B: N x m CSR (column k = measurement k), Sigma: m variances (already
 * multiplied by variance_scaling, src/linear_operator/measured_operator.cc:19). */
typedef struct {
  int dim;   /* 2 or 3 */
  int n[3];  /* cells per direction (nx, ny, nz); unknowns = prod(n_d - 1) */
  double kappa_sq;
  const int64_t* A_rowptr;
  const int64_t* A_col;
  const double* A_val;
  int m;
  const int64_t* B_rowptr;
  const int64_t* B_col;
  const double* B_val;
  const double* sigma;
  const double* diffusion;       /* NULL: a = 1 */
  const double* kappa_sq_field;  /* NULL: kappa_sq everywhere */
} mgmc_problem;

/* MultigridParameters (src/auxilliary/parameters.hh:150-179) plus the
 * MI355X-side choices (precision, number of chains, RNG seed, stream). */
typedef struct {
  int nlevel;
  int smoother;       /* mgmc_smoother */
  int npresmooth;
  int npostsmooth;
  int ncoarsesmooth;  /* SSOR coarse sampler only */
  double omega;
  int cycle;          /* 1 = V, 2 = W */
  double coarse_scaling;
  int coarse_solver;  /* mgmc_coarse (sampler); the preconditioner is always Cholesky */
  int precision;      /* mgmc_precision of the cycle (sweeps, transfers) */
  int nchains;        /* K independent chains batched in every launch */
  uint64_t seed;      /* Philox4x32-10 key */
  void* stream;       /* hipStream_t, NULL = a stream owned by the handle */
} mgmc_params;

/* Builds the hierarchy: replaces the constructors of MultigridMCSampler
 * (src/sampler/multigridmc_sampler.cc:8-106) and MultigridPreconditioner
 * (src/preconditioner/multigrid_preconditioner.cc:8-71). */
int mgmc_create(const mgmc_problem* problem, const mgmc_params* params, mgmc_handle** out);
int mgmc_destroy(mgmc_handle* h);
const char* mgmc_last_error(void);

/* Unknowns on `level` (Lattice::Nvertex, src/lattice/lattice.hh:126-131). */
int mgmc_ndof(const mgmc_handle* h, int level, int64_t* n);

/* Sampler::apply(f, x) for all chains (src/sampler/sampler.hh:48-49,
 * src/sampler/multigridmc_sampler.cc:139-145): one MGMC cycle per chain.
 * f: N values shared by all chains (f_per_chain = 0) or nchains x N.
 * x: nchains x N, chain state on entry, new sample on exit. */
int mgmc_sample(mgmc_handle* h, const double* f, int f_per_chain, double* x);

/* Resident sampling (the loop of src/measurements/sampling_time.cc:30-41):
 * the fine-level right-hand side and chain states stay in device memory.
 * mgmc_set_rhs replaces Sampler::fix_rhs (src/sampler/sampler.hh:64). */
int mgmc_set_rhs(mgmc_handle* h, const double* f, int f_per_chain);
int mgmc_set_state(mgmc_handle* h, const double* x);
int mgmc_get_state(mgmc_handle* h, double* x);
/* Observables: values of x at `nprobe` fine vertices (reference order). */
int mgmc_set_probes(mgmc_handle* h, int nprobe, const int64_t* vertices);
/* Observables as sparse vectors: z_p = sum_e weights[e] x[vertices[e]] for
 * e in [rowptr[p], rowptr[p+1]) (reference order; weights NULL = all 1).
 * Replaces the sample_vector.dot(x) of measure_sampling_time /
 * measure_convergence / measure_mean_squared_error
 * (src/measurements/sampling_time.cc:40, convergence.cc:57,
 * mean_squared_error.cc:45); a probe is one entry of weight 1. */
int mgmc_set_observables(mgmc_handle* h, int nobs, const int64_t* rowptr, const int64_t* vertices,
                         const double* weights);
/* Runs nsamples cycles on the device; if obs != NULL writes
 * obs[s][chain][probe] (host) after each cycle. use_graph = 1 replays a
 * captured hipGraph of the cycle. */
int mgmc_run(mgmc_handle* h, int nsamples, double* obs, int use_graph);
/* Same, with the observable buffer in device memory (nsamples x nchains x nprobe). */
int mgmc_run_dev(mgmc_handle* h, int nsamples, double* obs_dev, int use_graph);

/* ---- Batches of independent instances (no reference counterpart: the
 * reference runs one sampler at a time; these are for callers that sample
 * many problems of different sizes at once) --------------------------------
 * A batch runs the MGMC cycles of several handles (lattices of any sizes,
 * any chain counts, FP32 or FP64, each with its own measurements and seed)
 * together: at every point of the V-cycle the launches all instances make of
 * one kernel are merged into a single launch whose blocks are the instances'
 * blocks side by side, and one cycle of the whole batch is one captured
 * hipGraph. Every instance's samples equal those of its own mgmc_run_dev.
 * Requirements: handles on one device, multicolour order, constant-
 * coefficient stencils, V-cycles (MGMC_ERR_UNSUPPORTED otherwise). The
 * handles must outlive the batch and are used by it between calls. */
typedef struct mgmc_batch mgmc_batch;
int mgmc_batch_create(int n, mgmc_handle* const* handles, mgmc_batch** out);
int mgmc_batch_destroy(mgmc_batch* b);
/* nsamples cycles of every instance; obs_dev[i] (may be NULL, or the array
 * NULL) receives instance i's observables as mgmc_run_dev does. Work is
 * queued on the batch's stream; later work on the handles' streams waits
 * for it. */
int mgmc_batch_run_dev(mgmc_batch* b, int nsamples, double* const* obs_dev, int use_graph);
/* Persistent merged launches: at most workgroups_per_cu workgroups per CU
 * per launch, each looping over the blocks of all instances (0, the
 * default: one workgroup per block). */
int mgmc_batch_set_persistent(mgmc_batch* b, int workgroups_per_cu);
/* Launches per batch cycle after merging, and what the instances issue on
 * their own (either pointer may be NULL). */
int mgmc_batch_stats(const mgmc_batch* b, int64_t* launches_per_cycle, int64_t* separate_launches_per_cycle);
/* Pointwise posterior statistics (posterior_statistics,
 * src/measurements/posterior_statistics.cc:4-55): runs `nsamples` cycles from
 * the current state and returns the mean and variance of every fine vertex
 * over all samples and chains (reference order, N values each; either
 * pointer may be NULL). */
int mgmc_moments(mgmc_handle* h, int nsamples, double* mean, double* variance);
/* Reset the sample counter that, with the seed, keys the Philox noise. */
int mgmc_set_counter(mgmc_handle* h, uint64_t sample_counter);

/* Visiting order of every SOR sweep of the handle (default multicolour).
 * MGMC_ORDER_LEXICOGRAPHIC is the reference's own order: vertices in
 * increasing index forward and decreasing index backward
 * (src/smoother/sor_smoother.cc:56-77, src/sampler/sor_sampler.cc:37-59),
 * run as wavefronts of independent vertices, with the SMW correction of
 * that order (src/smoother/sor_smoother.cc:17-51); cycles and solves then
 * follow the reference's arithmetic and per-cycle residual reduction, at
 * many more launches per sweep. MGMC_ORDER_MULTICOLOUR (red-black / 2^d
 * colours) is the fast default; both are exact Gibbs samplers of the same
 * posterior. */
int mgmc_set_ordering(mgmc_handle* h, int ordering);

/* MultigridPreconditioner::apply (src/preconditioner/multigrid_preconditioner.cc:104-109):
 * one cycle from a zero initial guess, b, x: nchains x N. */
int mg_precondition(mgmc_handle* h, const double* b, double* x);

/* LoopSolver::apply (src/solver/loop_solver.cc:9-54): x = 0; repeat
 * r = Q x - b (FP64), stop if ||r||/||r0|| < rtol and ||r|| < atol, else
 * x -= MG(r). Chain 0 only. hist (maxiter entries, may be NULL) receives
 * ||r_k||; *niter the converged iteration or -1. The cycle runs in the
 * handle's precision, residuals and the update in FP64 (iterative refinement). */
int mg_solve(mgmc_handle* h, const double* b, double* x, double rtol, double atol, int maxiter, double* hist,
             int* niter);

/* mg_solve with b and x in device memory (nchains x N, reference order),
 * e.g. torch tensors; returns when x is written. */
int mg_solve_dev(mgmc_handle* h, const double* b_dev, double* x_dev, double rtol, double atol, int maxiter,
                 double* hist, int* niter);

/* The measurement operator on device vectors in reference order
 * (MeasuredOperator's B, src/linear_operator/measured_operator.cc:9-49):
 * z_dev (nchains x m) = B^T x_dev (nchains x N), and y_dev (nchains x N) =
 * B w_dev (nchains x m). Both return when the output is written. */
int mgmc_measurement_dot_dev(mgmc_handle* h, const double* x_dev, double* z_dev);
int mgmc_measurement_apply_dev(mgmc_handle* h, const double* w_dev, double* y_dev);

/* LinearOperator::apply (src/linear_operator/linear_operator.hh:78-88) on
 * level 0 in FP64, all chains: y = Q x. */
int mgmc_apply_operator(mgmc_handle* h, const double* x, double* y);

/* ---- component entry points (test hooks; same kernels as the cycle) ---- */

/* One SOR sweep on `level` in the handle's precision, all chains:
 * forward = 1 / backward = 0; stochastic = 1 adds the Gibbs noise of
 * SORSampler::apply (src/sampler/sor_sampler.cc:37-59), 0 is the smoother
 * SORSmoother::apply with nsmooth = 1 (src/smoother/sor_smoother.cc:41-53).
 * xi (nchains x N_level) and xi_lr (nchains x m) inject the normals; when
 * NULL they come from the Philox stream. b, x: nchains x N_level. */
int mgmc_sweep(mgmc_handle* h, int level, int forward, int stochastic, const double* b, double* x,
               const double* xi, const double* xi_lr);
/* fc = R (f - Q x) on level -> level+1 (multigridmc_sampler.cc:125-128). */
int mgmc_residual_restrict(mgmc_handle* h, int level, const double* f, const double* x, double* fc);
/* x += alpha P xc (src/intergrid/intergrid_operator.hh:106-120). */
int mgmc_prolongate_add(mgmc_handle* h, int level, double alpha, const double* xc, double* x);

/* ---- level components with the reference's semantics ----
 * Vectors are nchains x N_level (reference order, FP64); the sweeps run in
 * the handle's precision and ordering.
 * mgmc_smoother_apply: Smoother::apply(b, x) (src/smoother/smoother.hh:29) of
 * SORSmoother (smoother = MGMC_SOR, direction `forward`;
 * src/smoother/sor_smoother.cc:41-77: nsmooth times nsmooth sparse sweeps,
 * each followed by the SMW correction) or SSORSmoother (MGMC_SSOR;
 * src/smoother/ssor_smoother.cc:9-15: nsmooth times a forward and a backward
 * SORSmoother with nsmooth = 1). */
int mgmc_smoother_apply(mgmc_handle* h, int level, int smoother, int forward, int nsmooth, const double* b,
                        double* x);
/* Sampler::apply(f, x) of SORSampler / SSORSampler on `level`
 * (src/sampler/sor_sampler.cc:37-59, src/sampler/ssor_sampler.cc:9-15):
 * nsmooth Gibbs sweeps (SSOR: forward and backward pairs) with fresh Philox
 * noise; every call advances the handle's sample counter, so consecutive
 * calls draw independent noise. */
int mgmc_level_sample(mgmc_handle* h, int level, int smoother, int forward, int nsmooth, const double* f,
                      double* x);
/* IntergridOperator::restrict (src/intergrid/intergrid_operator.hh:74-88):
 * xc = P^T x from `level` to level + 1. */
int mgmc_restrict(mgmc_handle* h, int level, const double* x, double* xc);
/* Coarsest level: x = L^{-T}(xi + L^{-1} f) (src/sampler/cholesky_sampler.hh:52-66)
 * with xi injected (nchains x Nc) or from Philox; xi_mode = 0: deterministic
 * x = Q_c^{-1} f (src/solver/cholesky_solver.cc:30-40), 1: Philox, 2: injected. */
int mgmc_coarse(mgmc_handle* h, const double* f, double* x, int xi_mode, const double* xi);
/* Philox4x32-10 normals exactly as the sweeps draw them: for every chain and
 * fine vertex (reference order) the normal of sweep `tag` at the current
 * sample counter. out: nchains x N. */
int mgmc_normals(mgmc_handle* h, uint32_t tag, double* out);
/* ---- host-side operator setup for C/C++ callers (no device work) ---- */

/* Ball-average measurement vector of MeasuredOperator::measurement_vector
 * (src/linear_operator/measured_operator.cc:69-170): top-hat average over the
 * ball (x0, radius) tested against the multilinear hat functions, 2^d-point
 * Gauss quadrature per cell; radius < 1e-12 gives the nearest vertex. Writes
 * ascending reference vertex indices and values; *nnz receives the length
 * (MGMC_ERR_ARG if it exceeds `capacity`). n holds dim cell counts. */
int mgmc_measurement_vector(int dim, const int* n, const double* x0, double radius, int64_t capacity, int64_t* idx,
                            double* val, int64_t* nnz);
/* The measurement vector with a choice of weight: profile 0 is the top-hat
 * of mgmc_measurement_vector; profile 1 weights the ball average with
 * exp(-xi^2 / (2 width^2)), xi = |x - x0| / radius < 1, normalised to
 * integral one over the ball (a Gaussian-weighted ball average; no
 * counterpart in the reference, whose weight f_meas is 1,
 * src/linear_operator/measured_operator.hh:65). Same quadrature and output
 * as mgmc_measurement_vector. */
int mgmc_measurement_vector_weighted(int dim, const int* n, const double* x0, double radius, int profile,
                                     double width, int64_t capacity, int64_t* idx, double* val, int64_t* nnz);
/* Constant 3^d stencil of the Q1 finite-element prior -Delta + kappa^2
 * (src/linear_operator/shiftedlaplace_fem_operator.cc:9-141), index
 * (dz+1)*9 + (dy+1)*3 + (dx+1) (3D) or (dy+1)*3 + (dx+1) (2D); building A as a
 * CSR matrix from it and passing it to mgmc_create gives the FEM prior. */
int mgmc_fem_stencil(int dim, const int* n, double kappa_sq, double* stencil);

/* Host-side set-up only (no device work): the stencil of every vertex of
 * `level` of the hierarchy of `problem` (its A, or its coefficient fields;
 * measurements are ignored) with nlevel levels: out (N_level x 27, reference
 * vertex order) receives the coupling to each neighbour at index
 * (dz+1)*9 + (dy+1)*3 + (dx+1), zero where the neighbour is not interior.
 * For variable coefficients the coarse stencils are the Galerkin products
 * R A R^T (src/linear_operator/linear_operator.cc:10-23) formed per vertex. */
int mgmc_level_stencils(const mgmc_problem* problem, int nlevel, int level, double* out);

/* Adjoint contraction for parameter gradients (no counterpart in the
 * reference): for every chain c, out[2c] = lam_c . mu_c and
 * out[2c+1] = lam_c^T S mu_c over the fine interior vertices, S a constant
 * stencil (27 entries in 3D, 9 in 2D, index as mgmc_level_stencil) applied
 * with Dirichlet boundaries: the derivative of the prior operator with
 * respect to a coefficient (e.g. the unit-diffusion stencil for a diffusion
 * scale). lam, mu: nchains x N. */
int mgmc_contract(mgmc_handle* h, const double* stencil, const double* lam, const double* mu, double* out);

/* mgmc_contract without allocation. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). lam_dev, mu_dev:
 * device arrays of nchains x N FP64 in reference vertex order; out_dev:
 * device array of 2 x nchains receiving what mgmc_contract writes to `out`.
 * The partial sums use a fixed device array of 65,536 doubles (512 KiB per
 * GPU, held until the process ends); nchains x 2 x min(256, ceil(N / 2048))
 * must fit in it, else MGMC_ERR_ARG. Work is queued on the handle's stream
 * without synchronising; calls that overlap on one GPU race for the array. */
int mgmc_contract_noalloc(mgmc_handle* h, const double* stencil, const double* lam_dev, const double* mu_dev,
                          double* out_dev);

/* Export the level's stencil as used by the kernels (27 or 9 entries, x fastest,
 * offsets -1..1 per direction) and the colour count (2 = red-black). */
int mgmc_level_stencil(const mgmc_handle* h, int level, double* stencil, int* ncolours);

/* Launch-shape tuning (no reference counterpart). Keys: "fused_sweep" (0/1:
 * sweep 3D star levels with the single-pass red-black kernel), "rb_threads"
 * (threads per workgroup of that kernel, 64..1024; default 512), "rb_blocks"
 * (workgroups it and the fused residual-restriction aim for, default 768;
 * more z-segments when fewer tiles x chains exist), "rb_balanced" (0/1,
 * default 1: split the fused sweep's chains x tiles x planes evenly over one
 * round of resident workgroups instead of z-segments of rb_blocks),
 * "fold_prolongation" (0/1: add the coarse correction inside the first
 * post-smoothing sweep), "fused_residual" (0/1: residual + restriction in one
 * pass on 3D star levels), "rr_threads" (its threads per workgroup,
 * 64..1024), "plane_sweep" (0/1: sweep 3D box levels in two plane-parity
 * launches), "plane_sweep_min" (fewest plane x chain workgroups for that),
 * "plane_sweep_stage3" (0/1: planes k-1, k, k+1 in LDS), "coarse_mfma" (0/1:
 * coarsest-level products on the FP64 matrix cores), "rb_min_planes" /
 * "rr_min_planes" (fewest planes per z-segment of the fused level sweep /
 * residual-restriction), "lr_dot_chains" (1, 2, 4:
 * chains per workgroup of the low-rank dot B^T x), "lr_update_chains" (2, 4,
 * 8, 16: chains per thread of the low-rank update, with more than 8 chains),
 * "lr_carry" (0/1, default 1: the residual's B^T x taken from the SMW step
 * just before it, B^T (y + W u) = B^T y + (B^T W) u, instead of a second
 * gather over x), "lr_split" (>= 64; entries of B per workgroup of the
 * low-rank dot, whose columns are split into up to 8 parts: default 128 below
 * 8 chains, else 512). Results do not depend on these settings beyond
 * rounding. */
int mgmc_set_tuning(mgmc_handle* h, const char* key, int value);

/* Autotuner (no reference counterpart): coordinate search over the tuning
 * keys "rb_chains", "rb_balanced", "rb_threads", "rb_blocks", "rb_min_planes",
 * "rr_threads", "rr_min_planes" and "plane_sweep_stage3", which change launch shapes but no
 * computed value; every candidate is timed
 * over `ncycles` sampling cycles on the handle's stream and the fastest
 * setting is kept. The chains' state and the sample counter are restored,
 * so the sample path after tuning equals the untuned one bit for bit.
 * best_ms (may be NULL) receives the best time per cycle. */
int mgmc_autotune(mgmc_handle* h, int ncycles, double* best_ms);
/* Current value of one of the autotuned keys, or of "lr_carry" / "lr_split". */
int mgmc_get_tuning(const mgmc_handle* h, const char* key, int* value);

/* Profiling aid: with enable = 1 every level-0 sweep launch (one fused
 * red-black sweep, or one colour phase) is bracketed by HIP events on the
 * handle's stream (runs then launch eagerly).
 * mgmc_profile_result returns the mean duration of the launches recorded
 * since the last call and their number. No reference counterpart. */
int mgmc_profile(mgmc_handle* h, int enable);
int mgmc_profile_result(mgmc_handle* h, double* avg_ms, int64_t* count);

/* ---- Full-problem Cholesky samplers (host) ------------------------------
 * SparseCholeskySampler, LowRankCholeskySampler and DenseCholeskySampler
 * (src/sampler/cholesky_sampler.cc:9-129, cholesky_sampler.hh:52-86): exact
 * independent samples x ~ N(Q^{-1} f, Q^{-1}) of the problem's posterior
 * Q = A + B Sigma^{-1} B^T (or prior, m = 0), the baselines of the MGMC
 * sampler. Host code only (no GPU is used): nested-dissection ordering and a
 * supernodal multifrontal factorisation on `nthreads` host threads (0 = all
 * hardware threads), chains solved in blocks on those threads.
 * "sparse" factorises Q, "lowrank" factorises A and applies the reference's
 * QR-based correction for B Sigma^{-1} B^T, "dense" factorises Q without
 * reordering (at most 16384 unknowns). Replaces the constructors
 * (cholesky_sampler.cc:9-23, :26-38, :41-129). */
typedef struct mgmc_cholesky mgmc_cholesky;
enum mgmc_factorisation { MGMC_FACTOR_SPARSE = 0, MGMC_FACTOR_LOWRANK = 1, MGMC_FACTOR_DENSE = 2 };
int mgmc_cholesky_create(const mgmc_problem* problem, int factorisation, int nchains, uint64_t seed, int nthreads,
                         mgmc_cholesky** out);
int mgmc_cholesky_destroy(mgmc_cholesky* c);
/* Unknowns, entries of the factor L (the reference's nnz(L) print-out,
 * cholesky_wrapper.cc:85-91), supernodes and set-up seconds (any may be NULL). */
int mgmc_cholesky_info(const mgmc_cholesky* c, int64_t* n, int64_t* nnz_L, int64_t* nsupernodes, double* seconds);
/* Number of chains of the following mgmc_cholesky_sample calls (one
 * factorisation serves any number; the reference deep-copies its sampler per
 * thread instead, cholesky_sampler.hh:103-110). */
int mgmc_cholesky_set_nchains(mgmc_cholesky* c, int nchains);
/* CholeskySampler::fix_rhs / unfix_rhs (cholesky_sampler.hh:75-90). */
int mgmc_cholesky_fix_rhs(mgmc_cholesky* c, const double* f);
int mgmc_cholesky_unfix_rhs(mgmc_cholesky* c);
/* CholeskySampler::apply (cholesky_sampler.hh:52-66) for every chain: x
 * (nchains x N, reference order) receives independent samples. f: N values
 * shared by all chains (f_per_chain = 0) or nchains x N; ignored after
 * fix_rhs. xi: NULL draws the normals from Philox4x32-10 (one sample counter
 * per call); else nchains x N given normals, consumed in the factor's order
 * (a deterministic map xi -> x whose covariance is Q^{-1}; xi = 0 gives
 * the mean Q^{-1} f). */
int mgmc_cholesky_sample(mgmc_cholesky* c, const double* f, int f_per_chain, const double* xi, double* x);

#ifdef __cplusplus
}
#endif
#endif /* MGMC_AMD_H */
