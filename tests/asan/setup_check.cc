// Host set-up code of the library (setup.cc, operators.cc) under
// AddressSanitizer and UndefinedBehaviorSanitizer: hierarchies for 2D and 3D
// lattices whose x extent is and is not a multiple of 8 (the padded rows then
// do or do not share boundary slots), with ball measurements, the FD prior
// and a caller-supplied CSR prior (the Q1 FEM stencil), the multicolour SMW
// factors and the lexicographic capacitance matrices. Any out-of-bounds or
// undefined access aborts; every result is computed twice with heap churn in
// between and must repeat bit for bit. Prints "setup_check ok".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../multigridmc_amd/csrc/setup.h"

using namespace mgmc;

struct Case {
  int dim, n[3], m, nlevel;
  bool fem;
  bool variable = false;  // cell-wise diffusion and vertex-wise kappa^2 fields
};

static std::vector<double> run(const Case& c, double omega) {
  std::mt19937_64 rng(1234 + c.n[0] * 7 + c.m);
  std::uniform_real_distribution<double> U(0.2, 0.8);
  const Geom g = Geom::make(c.dim, c.n);
  // B as CSR rows from ball measurement vectors
  std::vector<std::vector<std::pair<int64_t, double>>> rows(g.N);
  std::vector<double> sigma(c.m);
  for (int k = 0; k < c.m; ++k) {
    double x0[3] = {U(rng), U(rng), U(rng)};
    std::vector<int64_t> idx;
    std::vector<double> val;
    measurement_vector(c.dim, c.n, x0, 0.15, idx, val);
    for (size_t e = 0; e < idx.size(); ++e) rows[idx[e]].push_back({k, val[e]});
    sigma[k] = 1e-3 * (1.0 + U(rng));
  }
  std::vector<int64_t> brp(1, 0), bci;
  std::vector<double> bv;
  for (auto& r : rows) {
    for (auto& e : r) {
      bci.push_back(e.first);
      bv.push_back(e.second);
    }
    brp.push_back((int64_t)bci.size());
  }
  // optional CSR prior from the FEM stencil
  std::vector<int64_t> arp(1, 0), aci;
  std::vector<double> av;
  if (c.fem) {
    double st[27];
    fem_stencil(c.dim, c.n, 10.0, st);
    const int zr = c.dim == 3 ? 1 : 0;
    for (int64_t r = 0; r < g.N; ++r) {
      int e[3];
      g.euclid(r, e);
      for (int dz = -zr; dz <= zr; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int b0 = e[0] + dx, b1 = e[1] + dy, b2 = e[2] + dz;
            if (b0 < 1 || b0 > c.n[0] - 1 || b1 < 1 || b1 > c.n[1] - 1) continue;
            if (c.dim == 3 && (b2 < 1 || b2 > c.n[2] - 1)) continue;
            const int64_t col = ((int64_t)(c.dim == 3 ? b2 - 1 : 0) * (c.n[1] - 1) + (b1 - 1)) * (c.n[0] - 1) + (b0 - 1);
            aci.push_back(col);
            av.push_back(st[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)]);
          }
      arp.push_back((int64_t)aci.size());
    }
  }
  std::vector<double> a_cell, kap;
  if (c.variable) {
    const int64_t ncell = (int64_t)c.n[0] * c.n[1] * (c.dim == 3 ? c.n[2] : 1);
    for (int64_t q = 0; q < ncell; ++q) a_cell.push_back(1.0 + 9.0 * U(rng));
    for (int64_t q = 0; q < g.N; ++q) kap.push_back(5.0 + 10.0 * U(rng));
  }
  HostProblem p{c.dim, {c.n[0], c.n[1], c.n[2]}, 10.0, c.fem ? arp.data() : nullptr, c.fem ? aci.data() : nullptr,
                c.fem ? av.data() : nullptr, c.m, brp.data(), bci.data(), bv.data(), sigma.data(),
                c.variable ? a_cell.data() : nullptr, c.variable ? kap.data() : nullptr};
  HostHierarchy H = build_hierarchy(p, c.nlevel, omega, true);
  std::vector<double> out;
  for (auto& L : H.levels) {
    out.insert(out.end(), L.st, L.st + 27);
    out.insert(out.end(), L.vst.begin(), L.vst.end());
    for (int d = 0; d < 2; ++d) {
      out.insert(out.end(), L.Sinv[d].begin(), L.Sinv[d].end());
      out.insert(out.end(), L.W[d].val.begin(), L.W[d].val.end());
    }
    if (c.m > 0) {
      std::vector<double> S[2];
      lex_capacitance(L, H.omega, H.sigma, S);
      for (int d = 0; d < 2; ++d) out.insert(out.end(), S[d].begin(), S[d].end());
    }
  }
  out.insert(out.end(), H.coarse.LinvT.begin(), H.coarse.LinvT.end());
  return out;
}

int main() {
  const Case cases[] = {
      {3, {16, 16, 16}, 3, 3, false}, {3, {12, 16, 8}, 3, 2, false}, {2, {16, 16, 1}, 2, 3, false},
      {2, {24, 20, 1}, 3, 2, false},  {3, {16, 16, 16}, 2, 3, true},  {2, {32, 32, 1}, 4, 3, true},
      {3, {16, 16, 16}, 3, 3, false, true}, {3, {12, 16, 8}, 2, 2, false, true}, {2, {32, 24, 1}, 3, 3, false, true},
  };
  for (const Case& c : cases)
    for (double omega : {1.0, 0.8}) {
      std::vector<double> a = run(c, omega);
      // heap churn: freed blocks full of non-zero bytes before the repeat
      for (int k = 0; k < 64; ++k) {
        std::vector<double> junk(4096 + 97 * k, 1e300);
        junk[k] = -1.0;
      }
      std::vector<double> b = run(c, omega);
      if (a.size() != b.size() || std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) != 0) {
        std::printf("setup_check: results differ between repetitions (dim %d, n %d x %d x %d)\n", c.dim, c.n[0], c.n[1],
                    c.n[2]);
        return 1;
      }
      for (double v : a)
        if (!std::isfinite(v)) {
          std::printf("setup_check: non-finite value\n");
          return 1;
        }
    }
  std::printf("setup_check ok\n");
  return 0;
}
