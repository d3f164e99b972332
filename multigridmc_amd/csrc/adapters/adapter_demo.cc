// Example use of the adapter classes with the reference's interface shapes
// (the calls src/driver_mg.cc and src/measurements/sampling_time.cc make on
// the reference's classes): a 2D 32 x 32 posterior with three ball
// measurements, solved by the loop solver with the multigrid preconditioner
// and sampled by MGMC chains through Sampler::apply, fix_rhs and deep_copy;
// then the level components: SOR / SSOR smoothers, SOR / SSOR samplers and
// the intergrid operator. The vectors go to OUTDIR as raw FP64 files for
// checking; the key of each sampler is printed.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "mgmc_adapters.hh"

using namespace mgmc_amd;

static void check_status(int status) {
  if (status != MGMC_OK) throw Error(status, mgmc_last_error());
}

static void save(const std::string& dir, const char* name, const std::vector<double>& v) {
  const std::string path = dir + "/" + name;
  FILE* fh = std::fopen(path.c_str(), "wb");
  if (!fh || std::fwrite(v.data(), sizeof(double), v.size(), fh) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(1);
  }
  std::fclose(fh);
}

int main(int argc, char* argv[]) {
  if (argc != 2) {
    std::printf("Usage: %s OUTDIR\n", argv[0]);
    return 1;
  }
  const std::string out = argv[1];
  try {
    Lattice lat;
    lat.dim = 2;
    lat.n[0] = lat.n[1] = 32;
    const std::vector<double> centres = {0.3, 0.4, 0.7, 0.55, 0.45, 0.75};
    const std::vector<double> variance = {1e-3, 2e-3, 1.5e-3};
    auto prior = LinearOperator::shifted_laplace_fd(lat, 10.0);
    auto op = LinearOperator::measured(prior, centres, 0.1, variance);
    const size_t N = (size_t)lat.Nvertex();
    MultigridParameters mp;
    mp.nlevel = 4;

    // driver_mg: right-hand side from std::mt19937_64(1482817), loop solver
    std::vector<double> b(N), x(N, 0.0), xp(N, 0.0);
    std::mt19937_64 gen(1482817);
    std::normal_distribution<double> normal(0.0, 1.0);
    for (auto& v : b) v = normal(gen);
    auto pre = std::make_shared<HipMultigridPreconditioner>(op, mp);
    pre->apply(b, xp);
    IterativeSolverParameters sp;
    sp.rtol = 1e-11;
    sp.atol = 1e-9;
    sp.maxiter = 50;
    HipLoopSolver solver(op, pre, sp);
    solver.apply(b, x);
    save(out, "b.bin", b);
    save(out, "precondition.bin", xp);
    save(out, "solve.bin", x);
    save(out, "history.bin", solver.history());

    // sampling: f = B Sigma^{-1} y (Q times the posterior mean for a zero prior mean)
    const std::vector<double> y = {0.5, -1.0, 0.25};
    const std::vector<double> f = op->B_Sigma_inv(y);
    auto sampler = std::make_shared<HipMultigridMCSampler>(op, std::make_shared<Mt19937Generator>(42), mp);
    const int ns = 5;
    std::vector<double> chain(N, 0.0), path, path_fixed, path_copy;
    for (int s = 0; s < ns; ++s) {
      sampler->apply(f, chain);
      path.insert(path.end(), chain.begin(), chain.end());
    }
    // the same key again (a generator with the same seed), now with f fixed on the device
    std::shared_ptr<Sampler> same = sampler->deep_copy(std::make_shared<Mt19937Generator>(42));
    same->fix_rhs(f);
    std::vector<double> none;
    std::fill(chain.begin(), chain.end(), 0.0);
    for (int s = 0; s < ns; ++s) {
      same->apply(none, chain);
      path_fixed.insert(path_fixed.end(), chain.begin(), chain.end());
    }
    // an independent chain from another generator
    std::shared_ptr<Sampler> other = sampler->deep_copy(std::make_shared<Mt19937Generator>(7));
    std::fill(chain.begin(), chain.end(), 0.0);
    for (int s = 0; s < ns; ++s) {
      other->apply(f, chain);
      path_copy.insert(path_copy.end(), chain.begin(), chain.end());
    }
    save(out, "f.bin", f);
    save(out, "samples.bin", path);
    save(out, "samples_fixed.bin", path_fixed);
    save(out, "samples_copy.bin", path_copy);
    std::printf("key %llu\n", (unsigned long long)sampler->seed());
    std::printf("key_copy %llu\n",
                (unsigned long long)std::static_pointer_cast<HipMultigridMCSampler>(other)->seed());
    // a length mismatch is reported, not undefined behaviour
    std::vector<double> short_x(N - 1);
    try {
      sampler->apply(f, short_x);
      std::printf("length check missing\n");
      return 1;
    } catch (const Error& e) {
      std::printf("length error status %d\n", e.status);
    }
    // Smoother, SOR / SSOR samplers and the intergrid operator (omega 0.8)
    std::vector<double> x0(N);
    for (auto& v : x0) v = normal(gen);
    save(out, "x0.bin", x0);
    std::vector<double> xs = x0;
    HipSORSmoother(op, 0.8, 1, Direction::forward).apply(b, xs);
    save(out, "sor_fwd.bin", xs);
    xs = x0;
    HipSORSmoother(op, 0.8, 1, Direction::backward).apply(b, xs);
    save(out, "sor_bwd.bin", xs);
    xs = x0;
    HipSSORSmoother(op, 0.8, 2).apply(b, xs);
    save(out, "ssor2.bin", xs);
    // the reference's nsmooth = 2 on the prior: twice (two sparse sweeps, then the correction)
    xs = x0;
    HipSORSmoother(prior, 0.8, 2, Direction::forward).apply(b, xs);
    save(out, "sor2_prior.bin", xs);
    // SORSampler on the prior: the normals of its first two calls and their results
    HipSORSampler sors(prior, std::make_shared<Mt19937Generator>(11), 0.8, 1, Direction::forward);
    std::vector<double> xi(N), xi2(N);
    check_status(mgmc_normals(sors.handle(), 1, xi.data()));
    xs = x0;
    sors.apply(b, xs);
    save(out, "sor_sampler_x1.bin", xs);
    check_status(mgmc_normals(sors.handle(), 1, xi2.data()));
    sors.apply(b, xs);
    save(out, "sor_sampler_x2.bin", xs);
    save(out, "sor_sampler_xi1.bin", xi);
    save(out, "sor_sampler_xi2.bin", xi2);
    // SSORSampler on the posterior: a path of three samples
    HipSSORSampler ssors(op, std::make_shared<Mt19937Generator>(12), 1.0, 1);
    std::vector<double> ssor_path;
    std::fill(chain.begin(), chain.end(), 0.0);
    for (int s = 0; s < 3; ++s) {
      ssors.apply(b, chain);
      ssor_path.insert(ssor_path.end(), chain.begin(), chain.end());
    }
    save(out, "ssor_sampler.bin", ssor_path);
    std::printf("key_ssor %llu\n", (unsigned long long)ssors.seed());
    // IntergridOperator: x_c = P^T x0, then x0 + 0.7 P x_c
    HipIntergridOperator P(op);
    std::vector<double> xc((size_t)P.coarse_ndof());
    P.restrict(x0, xc);
    save(out, "restrict.bin", xc);
    xs = x0;
    P.prolongate_add(0.7, xc, xs);
    save(out, "prolongate.bin", xs);
  } catch (const Error& e) {
    std::printf("ERROR: %s\n", e.what());
    return 1;
  }
  std::printf("adapter_demo done\n");
  return 0;
}
