// Full-problem Cholesky samplers on the host: the reference's
// SparseCholeskySampler, LowRankCholeskySampler and DenseCholeskySampler
// (src/sampler/cholesky_sampler.cc:9-129, cholesky_sampler.hh:52-66), the
// baselines the MGMC sampler is compared with.
//
// Factorisation. The precision matrix is reordered by geometric nested
// dissection of the lattice (planes of vertices as separators, the halves
// first; a measurement ball that straddles a separator joins it), then
// factorised by a supernodal multifrontal Cholesky: elimination tree and
// postorder, column counts from row subtrees, supernodes (fundamental ones,
// small subtrees and chains merged where few stored zeros result), and per
// supernode one dense frontal matrix (original entries plus the children's
// update matrices), partially factorised by blocked right-looking kernels
// whose trailing update runs on host threads. This is
// the work CHOLMOD does for the reference (src/auxilliary/cholesky_wrapper.cc,
// which lets CHOLMOD choose an AMD ordering); an ordering changes the map
// from the normals to the sample, not the distribution N(Q^{-1} f, Q^{-1}).
// The dense factorisation is the same code with one supernode holding every
// column in the reference's vertex order.
//
// Sampling (CholeskySampler::apply). With P Q P^T = L L^T,
//   x = P^T L^{-T} (xi + L^{-1} P f),  xi ~ N(0, I),
// with L^{-1} P f cached by fix_rhs. The low-rank sampler factorises only the
// prior A (P A P^T = L L^T) and folds in the measurements as the reference
// does (cholesky_sampler.cc:41-129): V = L^{-1} P B = Q_V R (thin QR),
// Id - Lambda = Id - R (Sigma + V^T V)^{-1} R^T = W W^T (W lower), then
//   g  = L^{-1} P f + Q_V (W^T - Id) Q_V^T L^{-1} P f,
//   xi <- xi + g,  xi <- xi + Q_V (W - Id) Q_V^T xi,  x = P^T L^{-T} xi.
// Noise: Philox4x32-10 keyed by the seed, counter (position / 4 in the factor
// order, chain, tag, sample), four Box-Muller normals per call; chains are
// solved in blocks of up to eight right-hand sides on host threads.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/mgmc_amd.h"
#include "philox.h"
#include "setup.h"

namespace mgmc {

namespace {

constexpr uint32_t kCholeskyTag = 0x43u;  // Philox counter word of the Cholesky noise
constexpr int kRhsBlock = 8;              // chains solved together by one thread
constexpr int kNb = 64;                   // column block of the dense partial factorisation
constexpr int64_t kMaxCliqueEntries = 200000000;
constexpr int32_t kRelax = 32;            // elimination subtrees up to this size form one supernode

// Runs fn(i) for i in [0, count) on up to nthreads threads (dynamic scheduling).
void parallel_for(int nthreads, int64_t count, const std::function<void(int64_t)>& fn) {
  if (nthreads <= 1 || count <= 1) {
    for (int64_t i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next(0);
  auto work = [&]() {
    for (int64_t i = next++; i < count; i = next++) fn(i);
  };
  const int nt = (int)std::min<int64_t>(nthreads, count);
  std::vector<std::thread> ts;
  for (int t = 1; t < nt; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
}

// Symmetric matrix (both triangles) in CSR, rows sorted, duplicates merged.
struct Csr {
  int64_t n = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
};

Csr to_csr(std::vector<std::vector<std::pair<int32_t, double>>>& rows) {
  Csr A;
  A.n = (int64_t)rows.size();
  A.rp.assign(1, 0);
  for (auto& r : rows) {
    std::sort(r.begin(), r.end(), [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
      return a.first < b.first;
    });
    for (size_t e = 0; e < r.size(); ++e) {
      if (!A.ci.empty() && (int64_t)A.ci.size() > A.rp.back() && A.ci.back() == r[e].first) {
        A.v.back() += r[e].second;
      } else {
        A.ci.push_back(r[e].first);
        A.v.push_back(r[e].second);
      }
    }
    A.rp.push_back((int64_t)A.ci.size());
    std::vector<std::pair<int32_t, double>>().swap(r);
  }
  return A;
}

// Geometric nested dissection of the interior vertices: the middle plane of
// the longest extent is ordered after the two halves, recursively; boxes of
// at most kLeaf vertices (or too thin to split) are ordered lexicographically.
// `cliques` (the vertex sets of the measurement balls, coupled pairwise by
// B Sigma^{-1} B^T) keep the halves independent: a clique with vertices in
// both halves of a box joins that box's separator with all of its vertices
// in the box, so no entry of the matrix couples the two halves.
std::vector<int32_t> nested_dissection(const Geom& g, const std::vector<std::vector<int32_t>>& cliques) {
  constexpr int64_t kLeaf = 16;
  const int m[3] = {g.n[0] - 1, g.n[1] - 1, g.dim == 3 ? g.n[2] - 1 : 1};
  std::vector<int32_t> order;
  order.reserve((size_t)g.N);
  std::vector<char> done((size_t)g.N, 0);
  auto coord = [&](int32_t v, int* c) {
    c[0] = (int)(v % m[0]);
    c[1] = (int)((v / m[0]) % m[1]);
    c[2] = (int)(v / ((int64_t)m[0] * m[1]));
  };
  auto inside = [](const int* c, const int* lo, const int* hi) {
    return c[0] >= lo[0] && c[0] < hi[0] && c[1] >= lo[1] && c[1] < hi[1] && c[2] >= lo[2] && c[2] < hi[2];
  };
  auto append = [&](const int* lo, const int* hi, std::vector<int32_t>& out) {
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i) {
          const int32_t v = (int32_t)(i + (int64_t)m[0] * (j + (int64_t)m[1] * k));
          if (!done[v]) {
            done[v] = 1;
            out.push_back(v);
          }
        }
  };
  std::function<void(const int*, const int*, const std::vector<int32_t>&)> rec =
      [&](const int* lo, const int* hi, const std::vector<int32_t>& balls) {
        int64_t vol = 1;
        int d = 0;
        for (int e = 0; e < 3; ++e) {
          vol *= hi[e] - lo[e];
          if (hi[e] - lo[e] > hi[d] - lo[d]) d = e;
        }
        if (vol <= 0) return;
        if (vol <= kLeaf || hi[d] - lo[d] < 3) {
          append(lo, hi, order);
          return;
        }
        const int mid = lo[d] + (hi[d] - lo[d]) / 2;
        int a_hi[3] = {hi[0], hi[1], hi[2]}, b_lo[3] = {lo[0], lo[1], lo[2]};
        int s_lo[3] = {lo[0], lo[1], lo[2]}, s_hi[3] = {hi[0], hi[1], hi[2]};
        a_hi[d] = mid;
        b_lo[d] = mid + 1;
        s_lo[d] = mid;
        s_hi[d] = mid + 1;
        std::vector<int32_t> sep, balls_a, balls_b;
        append(s_lo, s_hi, sep);
        int c[3];
        for (int32_t q : balls) {
          bool in_a = false, in_b = false;
          for (int32_t v : cliques[q]) {
            if (done[v]) continue;
            coord(v, c);
            if (!inside(c, lo, hi)) continue;
            in_a = in_a || c[d] < mid;
            in_b = in_b || c[d] > mid;
          }
          if (in_a && in_b) {
            for (int32_t v : cliques[q]) {
              coord(v, c);
              if (!done[v] && inside(c, lo, hi)) {
                done[v] = 1;
                sep.push_back(v);
              }
            }
          } else if (in_a) {
            balls_a.push_back(q);
          } else if (in_b) {
            balls_b.push_back(q);
          }
        }
        rec(lo, a_hi, balls_a);
        rec(b_lo, hi, balls_b);
        order.insert(order.end(), sep.begin(), sep.end());
      };
  const int lo[3] = {0, 0, 0};
  std::vector<int32_t> all((size_t)cliques.size());
  for (size_t q = 0; q < cliques.size(); ++q) all[q] = (int32_t)q;
  rec(lo, m, all);
  if ((int64_t)order.size() != g.N) throw Error(MGMC_ERR_ARG, "internal error: nested dissection lost vertices");
  return order;
}

// ---- dense kernels on a frontal matrix F (m x m, column-major, lower part) ----

// F[r, q] -= sum_{p in [c0, c1)} F[r, p] F[q, p] for rows r >= q (and some
// entries above the diagonal inside 8 x 4 tiles, which are never read) and
// columns q in [c1, m). The panel F[c1:m, c0:c1] is packed in blocks of 8 rows.
void trailing_update(double* F, int64_t m, int64_t c0, int64_t c1, int nthreads) {
  const int64_t nbk = c1 - c0, rows = m - c1;
  if (rows <= 0 || nbk <= 0) return;
  const int64_t nrb = (rows + 7) / 8;
  std::vector<double> pack((size_t)(nrb * nbk * 8), 0.0);
  for (int64_t p = 0; p < nbk; ++p) {
    const double* col = F + (c0 + p) * m + c1;
    for (int64_t r = 0; r < rows; ++r) pack[(size_t)(((r >> 3) * nbk + p) * 8 + (r & 7))] = col[r];
  }
  const int64_t ncb = (rows + 3) / 4;  // column blocks of 4
  const double flops = (double)rows * (double)rows * (double)nbk;
  auto colblock = [&](int64_t cb) {
    const int64_t q0 = cb * 4;  // local column offset
    const int nq = (int)std::min<int64_t>(4, rows - q0);
    // B operand: the 4 columns' rows of the panel
    double b[kNb][4];
    for (int64_t p = 0; p < nbk; ++p)
      for (int t = 0; t < 4; ++t)
        b[p][t] = t < nq ? pack[(size_t)((((q0 + t) >> 3) * nbk + p) * 8 + ((q0 + t) & 7))] : 0.0;
    for (int64_t rb = q0 >> 3; rb < nrb; ++rb) {
      double acc[4][8] = {};
      const double* A = pack.data() + (size_t)(rb * nbk * 8);
      for (int64_t p = 0; p < nbk; ++p) {
        const double* a = A + p * 8;
        for (int t = 0; t < 4; ++t) {
          const double bt = b[p][t];
          for (int i = 0; i < 8; ++i) acc[t][i] += a[i] * bt;
        }
      }
      const int64_t r0 = rb * 8;
      const int nr = (int)std::min<int64_t>(8, rows - r0);
      for (int t = 0; t < nq; ++t) {
        double* col = F + (c1 + q0 + t) * m + c1 + r0;
        for (int i = 0; i < nr; ++i) col[i] -= acc[t][i];
      }
    }
  };
  parallel_for(flops > 4e6 ? nthreads : 1, ncb, colblock);
}

// Factorises the first k columns of F in place: F[:, :k] becomes the panel
// [L11; L21] and F[k:, k:] the update matrix F22 - L21 L21^T.
void partial_cholesky(double* F, int64_t m, int64_t k, int nthreads) {
  for (int64_t c0 = 0; c0 < k; c0 += kNb) {
    const int64_t c1 = std::min<int64_t>(c0 + kNb, k);
    for (int64_t j = c0; j < c1; ++j) {
      double* cj = F + j * m;
      const double d = cj[j];
      if (!(d > 0.0)) throw Error(MGMC_ERR_NOT_SPD, "the matrix is not positive definite (Cholesky pivot <= 0)");
      const double s = std::sqrt(d), inv = 1.0 / s;
      cj[j] = s;
      for (int64_t r = j + 1; r < m; ++r) cj[r] *= inv;
      for (int64_t q = j + 1; q < c1; ++q) {
        const double a = cj[q];
        double* cq = F + q * m;
        for (int64_t r = q; r < m; ++r) cq[r] -= a * cj[r];
      }
    }
    trailing_update(F, m, c0, c1, nthreads);
  }
}

}  // namespace

// Supernodal Cholesky factor of P Q P^T.
struct SupernodalFactor {
  int64_t n = 0;
  std::vector<int32_t> perm;   // factor position -> reference vertex
  std::vector<int32_t> iperm;  // reference vertex -> factor position
  int64_t ns = 0;
  std::vector<int32_t> sfirst;  // ns + 1: columns [sfirst[s], sfirst[s+1])
  std::vector<int64_t> rptr;    // ns + 1: row indices of s in ridx[rptr[s], rptr[s+1])
  std::vector<int32_t> ridx;    // own columns first, then the rows below, ascending
  std::vector<int32_t> sparent;
  std::vector<int64_t> lptr;  // ns + 1: panel of s (m_s x k_s, column-major) at lval[lptr[s]]
  std::vector<double> lval;
  int64_t nnz = 0;  // entries of L (lower triangle of the diagonal blocks and the rows below)

  int64_t ncol(int64_t s) const { return sfirst[s + 1] - sfirst[s]; }
  int64_t nrow(int64_t s) const { return rptr[s + 1] - rptr[s]; }

  // ordering + symbolic analysis of the pattern of A (reference order)
  void analyse(const Csr& A, const std::vector<int32_t>& order, bool dense) {
    n = A.n;
    if (dense) {
      perm.resize((size_t)n);
      for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
      iperm = perm;
      ns = 1;
      sfirst = {0, (int32_t)n};
      rptr = {0, n};
      ridx = perm;
      sparent = {-1};
      finish_layout();
      return;
    }
    std::vector<int32_t> ip((size_t)n);
    for (int64_t k = 0; k < n; ++k) ip[order[k]] = (int32_t)k;
    // elimination tree (Liu's algorithm with path compression)
    std::vector<int32_t> parent((size_t)n, -1), anc((size_t)n, -1);
    for (int64_t j = 0; j < n; ++j) {
      const int32_t r = order[j];
      for (int64_t e = A.rp[r]; e < A.rp[r + 1]; ++e) {
        int32_t i = ip[A.ci[e]];
        while (i != -1 && i < j) {
          const int32_t nx = anc[i];
          anc[i] = (int32_t)j;
          if (nx == -1) parent[i] = (int32_t)j;
          i = nx;
        }
      }
    }
    // postorder (children in increasing order)
    std::vector<int32_t> head((size_t)n, -1), next((size_t)n, -1), post;
    post.reserve((size_t)n);
    for (int64_t j = n - 1; j >= 0; --j)
      if (parent[j] != -1) {
        next[j] = head[parent[j]];
        head[parent[j]] = (int32_t)j;
      }
    std::vector<int32_t> stack;
    for (int64_t root = 0; root < n; ++root) {
      if (parent[root] != -1) continue;
      stack.push_back((int32_t)root);
      while (!stack.empty()) {
        const int32_t p = stack.back();
        const int32_t c = head[p];
        if (c == -1) {
          post.push_back(p);
          stack.pop_back();
        } else {
          head[p] = next[c];
          stack.push_back(c);
        }
      }
    }
    std::vector<int32_t> ipost((size_t)n);
    for (int64_t k = 0; k < n; ++k) ipost[post[k]] = (int32_t)k;
    perm.resize((size_t)n);
    iperm.resize((size_t)n);
    std::vector<int32_t> par((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
      perm[k] = order[post[k]];
      iperm[perm[k]] = (int32_t)k;
      par[k] = parent[post[k]] == -1 ? -1 : ipost[parent[post[k]]];
    }
    // column counts (with the diagonal) from the row subtrees
    std::vector<int64_t> cc((size_t)n, 1);
    std::vector<int32_t> mark((size_t)n, -1);
    for (int64_t i = 0; i < n; ++i) {
      mark[i] = (int32_t)i;
      const int32_t r = perm[i];
      for (int64_t e = A.rp[r]; e < A.rp[r + 1]; ++e) {
        int32_t k = iperm[A.ci[e]];
        if (k >= i) continue;
        while (mark[k] != i) {
          ++cc[k];
          mark[k] = (int32_t)i;
          k = par[k];
        }
      }
    }
    // supernodes: every maximal elimination subtree of at most kRelax
    // columns becomes one (relaxed) supernode, whose columns are contiguous in
    // the postorder; elsewhere fundamental supernodes (j joins j-1's when j is
    // its parent and struct(j-1) = {j-1} + struct(j)).
    std::vector<int32_t> desc((size_t)n, 1);
    for (int64_t j = 0; j < n; ++j)
      if (par[j] != -1) desc[par[j]] += desc[j];
    std::vector<char> relaxed((size_t)n, 0), block_start((size_t)n, 0);
    for (int64_t j = 0; j < n; ++j)
      if (desc[j] <= kRelax && (par[j] == -1 || desc[par[j]] > kRelax)) {
        const int64_t first = j - desc[j] + 1;
        block_start[first] = 1;
        for (int64_t c = first; c <= j; ++c) relaxed[c] = 1;
      }
    // Along a chain (j-1's parent is j) columns are merged while the stored
    // zeros stay few, as CHOLMOD's relaxed amalgamation does: the merged
    // columns f..j hold rows {f..j} + struct(j) \ {j}.
    sfirst.assign(1, 0);
    int64_t f = 0;
    double actual = (double)cc[0];
    for (int64_t j = 1; j < n; ++j) {
      bool start;
      if (relaxed[j]) {
        start = block_start[j] != 0;
      } else if (relaxed[j - 1] || par[j - 1] != j) {
        start = true;
      } else if (cc[j - 1] == cc[j] + 1) {
        start = false;  // fundamental
      } else {
        const double w = (double)(j - f + 1), rows = w + (double)cc[j] - 1.0;
        const double stored = w * rows - w * (w - 1.0) / 2.0;
        const double zfrac = 1.0 - (actual + (double)cc[j]) / stored;
        start = !(w <= 4 || (w <= 16 && zfrac < 0.8) || (w <= 48 && zfrac < 0.1) || zfrac < 0.05);
      }
      if (start) {
        sfirst.push_back((int32_t)j);
        f = j;
        actual = 0.0;
      }
      actual += (double)cc[j];
    }
    ns = (int64_t)sfirst.size();
    sfirst.push_back((int32_t)n);
    std::vector<int32_t> super_of((size_t)n);
    for (int64_t s = 0; s < ns; ++s)
      for (int32_t j = sfirst[s]; j < sfirst[s + 1]; ++j) super_of[j] = (int32_t)s;
    sparent.assign((size_t)ns, -1);
    std::vector<std::vector<int32_t>> children((size_t)ns);
    for (int64_t s = 0; s < ns; ++s) {
      const int32_t p = par[sfirst[s + 1] - 1];
      if (p != -1) {
        sparent[s] = super_of[p];
        children[sparent[s]].push_back((int32_t)s);
      }
    }
    // row structure per supernode
    rptr.assign(1, 0);
    ridx.clear();
    std::fill(mark.begin(), mark.end(), -1);
    std::vector<int32_t> below;
    for (int64_t s = 0; s < ns; ++s) {
      const int32_t f = sfirst[s], l = sfirst[s + 1] - 1;
      below.clear();
      for (int32_t j = f; j <= l; ++j) {
        const int32_t r = perm[j];
        for (int64_t e = A.rp[r]; e < A.rp[r + 1]; ++e) {
          const int32_t i = iperm[A.ci[e]];
          if (i > l && mark[i] != s) {
            mark[i] = (int32_t)s;
            below.push_back(i);
          }
        }
      }
      for (int32_t c : children[s])
        for (int64_t e = rptr[c] + ncol(c); e < rptr[c + 1]; ++e) {
          const int32_t i = ridx[e];
          if (i > l && mark[i] != s) {
            mark[i] = (int32_t)s;
            below.push_back(i);
          }
        }
      std::sort(below.begin(), below.end());
      for (int32_t j = f; j <= l; ++j) ridx.push_back(j);
      ridx.insert(ridx.end(), below.begin(), below.end());
      rptr.push_back((int64_t)ridx.size());
      if (relaxed[f] ? nrow(s) < cc[f] : nrow(s) != (l - f) + cc[l])
        throw Error(MGMC_ERR_ARG, "internal error: supernode structure disagrees with column counts");
    }
    finish_layout();
  }

  void finish_layout() {
    lptr.assign(1, 0);
    nnz = 0;
    for (int64_t s = 0; s < ns; ++s) {
      const int64_t m = nrow(s), k = ncol(s);
      lptr.push_back(lptr.back() + m * k);
      nnz += k * (k + 1) / 2 + (m - k) * k;
    }
  }

  // numeric multifrontal factorisation of the values of A (same pattern)
  void factorise(const Csr& A, int nthreads) {
    lval.assign((size_t)lptr.back(), 0.0);
    std::vector<int32_t> pos((size_t)n, -1);
    struct Update {
      int64_t s;
      std::vector<double> U;
    };
    std::vector<Update> stack;
    std::vector<int> nchild((size_t)ns, 0);
    for (int64_t s = 0; s < ns; ++s)
      if (sparent[s] != -1) ++nchild[sparent[s]];
    std::vector<double> F;
    for (int64_t s = 0; s < ns; ++s) {
      const int64_t m = nrow(s), k = ncol(s);
      const int32_t* idx = ridx.data() + rptr[s];
      for (int64_t a = 0; a < m; ++a) pos[idx[a]] = (int32_t)a;
      F.assign((size_t)(m * m), 0.0);
      for (int64_t q = 0; q < k; ++q) {
        const int32_t j = idx[q], r = perm[j];
        for (int64_t e = A.rp[r]; e < A.rp[r + 1]; ++e) {
          const int32_t i = iperm[A.ci[e]];
          if (i >= j) F[(size_t)(pos[i] + q * m)] += A.v[e];
        }
      }
      for (int c = 0; c < nchild[s]; ++c) {
        if (stack.empty() || sparent[stack.back().s] != s)
          throw Error(MGMC_ERR_ARG, "internal error: update matrices out of order");
        const Update& u = stack.back();
        const int64_t cs = u.s, mu = nrow(cs) - ncol(cs);
        const int32_t* cidx = ridx.data() + rptr[cs] + ncol(cs);
        for (int64_t b = 0; b < mu; ++b) {
          double* Fc = F.data() + (size_t)pos[cidx[b]] * m;
          const double* Uc = u.U.data() + (size_t)(b * mu);
          for (int64_t a = b; a < mu; ++a) Fc[pos[cidx[a]]] += Uc[a];
        }
        stack.pop_back();
      }
      partial_cholesky(F.data(), m, k, nthreads);
      std::memcpy(lval.data() + lptr[s], F.data(), (size_t)(m * k) * sizeof(double));
      if (m > k) {
        Update u;
        u.s = s;
        const int64_t mu = m - k;
        u.U.resize((size_t)(mu * mu));
        for (int64_t b = 0; b < mu; ++b)
          std::memcpy(u.U.data() + (size_t)(b * mu), F.data() + (size_t)((k + b) * m + k), (size_t)mu * sizeof(double));
        stack.push_back(std::move(u));
      }
      for (int64_t a = 0; a < m; ++a) pos[idx[a]] = -1;
    }
    if (!stack.empty()) throw Error(MGMC_ERR_ARG, "internal error: update matrices left over");
  }

  // X (n x r, row-major, factor order) <- L^{-1} X; R = r when fixed at compile time
  template <int R>
  void solve_L_t(double* X, int rr) const {
    const int r = R > 0 ? R : rr;
    std::vector<double> tmp;
    for (int64_t s = 0; s < ns; ++s) {
      const int64_t m = nrow(s), k = ncol(s);
      const double* L = lval.data() + lptr[s];
      double* xs = X + (size_t)sfirst[s] * r;
      for (int64_t j = 0; j < k; ++j) {
        const double* lj = L + j * m;
        double* xj = xs + j * r;
        const double inv = 1.0 / lj[j];
        for (int c = 0; c < r; ++c) xj[c] *= inv;
        for (int64_t i = j + 1; i < k; ++i) {
          const double l = lj[i];
          double* xi = xs + i * r;
          for (int c = 0; c < r; ++c) xi[c] -= l * xj[c];
        }
      }
      if (m == k) continue;
      tmp.assign((size_t)((m - k) * r), 0.0);
      for (int64_t j = 0; j < k; ++j) {
        const double* lj = L + j * m + k;
        const double* xj = xs + j * r;
        for (int64_t i = 0; i < m - k; ++i) {
          const double l = lj[i];
          double* t = tmp.data() + i * r;
          for (int c = 0; c < r; ++c) t[c] += l * xj[c];
        }
      }
      const int32_t* idx = ridx.data() + rptr[s] + k;
      for (int64_t i = 0; i < m - k; ++i) {
        double* xi = X + (size_t)idx[i] * r;
        const double* t = tmp.data() + i * r;
        for (int c = 0; c < r; ++c) xi[c] -= t[c];
      }
    }
  }

  // sum_i a[i] b[i] in eight independent partial sums (vectorisable)
  static double dot(const double* a, const double* b, int64_t n) {
    double p[8] = {};
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
      for (int e = 0; e < 8; ++e) p[e] += a[i + e] * b[i + e];
    double s = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    for (; i < n; ++i) s += a[i] * b[i];
    return s;
  }

  // X <- L^{-T} X
  template <int R>
  void solve_LT_t(double* X, int rr) const {
    const int r = R > 0 ? R : rr;
    std::vector<double> tmp, accv((size_t)r);
    double* acc = accv.data();
    for (int64_t s = ns - 1; s >= 0; --s) {
      const int64_t m = nrow(s), k = ncol(s);
      const double* L = lval.data() + lptr[s];
      double* xs = X + (size_t)sfirst[s] * r;
      if (m > k) {
        const int32_t* idx = ridx.data() + rptr[s] + k;
        tmp.resize((size_t)((m - k) * r));
        for (int64_t i = 0; i < m - k; ++i)
          for (int c = 0; c < r; ++c) tmp[(size_t)(i * r + c)] = X[(size_t)idx[i] * r + c];
        for (int64_t j = 0; j < k; ++j) {
          const double* lj = L + j * m + k;
          double* xj = xs + j * r;
          if (R == 1) {
            xj[0] -= dot(lj, tmp.data(), m - k);
            continue;
          }
          for (int c = 0; c < r; ++c) acc[c] = 0.0;
          for (int64_t i = 0; i < m - k; ++i) {
            const double l = lj[i];
            const double* t = tmp.data() + i * r;
            for (int c = 0; c < r; ++c) acc[c] += l * t[c];
          }
          for (int c = 0; c < r; ++c) xj[c] -= acc[c];
        }
      }
      for (int64_t j = k - 1; j >= 0; --j) {
        const double* lj = L + j * m;
        double* xj = xs + j * r;
        const double inv = 1.0 / lj[j];
        if (R == 1) {
          xj[0] = (xj[0] - dot(lj + j + 1, xs + j + 1, k - j - 1)) * inv;
          continue;
        }
        for (int c = 0; c < r; ++c) acc[c] = 0.0;
        for (int64_t i = j + 1; i < k; ++i) {
          const double l = lj[i];
          const double* xi = xs + i * r;
          for (int c = 0; c < r; ++c) acc[c] += l * xi[c];
        }
        for (int c = 0; c < r; ++c) xj[c] = (xj[c] - acc[c]) * inv;
      }
    }
  }

  void solve_L(double* X, int r) const {
    switch (r) {
      case 1: return solve_L_t<1>(X, r);
      case 2: return solve_L_t<2>(X, r);
      case 4: return solve_L_t<4>(X, r);
      case 8: return solve_L_t<8>(X, r);
      default: return solve_L_t<0>(X, r);
    }
  }

  void solve_LT(double* X, int r) const {
    switch (r) {
      case 1: return solve_LT_t<1>(X, r);
      case 2: return solve_LT_t<2>(X, r);
      case 4: return solve_LT_t<4>(X, r);
      case 8: return solve_LT_t<8>(X, r);
      default: return solve_LT_t<0>(X, r);
    }
  }
};

namespace {

// Householder QR of V (n x m, column-major): Q (n x m, orthonormal columns)
// and R = Q^T V (m x m, row-major, upper triangular).
void householder_qr(int64_t n, int m, std::vector<double> V, std::vector<double>& Q, std::vector<double>& R) {
  std::vector<std::vector<double>> vs((size_t)m);
  std::vector<double> beta((size_t)m, 0.0);
  for (int j = 0; j < m; ++j) {
    double* a = V.data() + (size_t)j * n;
    double nrm = 0.0;
    for (int64_t i = j; i < n; ++i) nrm += a[i] * a[i];
    nrm = std::sqrt(nrm);
    std::vector<double> v((size_t)(n - j));
    const double alpha = a[j] >= 0.0 ? -nrm : nrm;
    for (int64_t i = j; i < n; ++i) v[(size_t)(i - j)] = a[i];
    v[0] -= alpha;
    double vv = 0.0;
    for (double t : v) vv += t * t;
    beta[j] = vv > 0.0 ? 2.0 / vv : 0.0;
    for (int c = j; c < m; ++c) {
      double* col = V.data() + (size_t)c * n;
      double s = 0.0;
      for (int64_t i = j; i < n; ++i) s += v[(size_t)(i - j)] * col[i];
      s *= beta[j];
      for (int64_t i = j; i < n; ++i) col[i] -= s * v[(size_t)(i - j)];
    }
    vs[j].swap(v);
  }
  Q.assign((size_t)n * m, 0.0);
  for (int j = 0; j < m; ++j) Q[(size_t)j * n + j] = 1.0;
  for (int j = m - 1; j >= 0; --j)
    for (int c = 0; c < m; ++c) {
      double* col = Q.data() + (size_t)c * n;
      double s = 0.0;
      for (int64_t i = j; i < n; ++i) s += vs[j][(size_t)(i - j)] * col[i];
      s *= beta[j];
      for (int64_t i = j; i < n; ++i) col[i] -= s * vs[j][(size_t)(i - j)];
    }
  R.assign((size_t)m * m, 0.0);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r <= c; ++r) R[(size_t)r * m + c] = V[(size_t)c * n + r];
}

// Lower Cholesky factor of a dense SPD m x m matrix (row-major).
std::vector<double> small_cholesky(std::vector<double> A, int m, const char* what) {
  for (int j = 0; j < m; ++j) {
    double s = A[(size_t)j * m + j];
    for (int k = 0; k < j; ++k) s -= A[(size_t)j * m + k] * A[(size_t)j * m + k];
    if (!(s > 0.0)) throw Error(MGMC_ERR_NOT_SPD, std::string(what) + " is not positive definite");
    const double d = std::sqrt(s);
    A[(size_t)j * m + j] = d;
    for (int i = j + 1; i < m; ++i) {
      double t = A[(size_t)i * m + j];
      for (int k = 0; k < j; ++k) t -= A[(size_t)i * m + k] * A[(size_t)j * m + k];
      A[(size_t)i * m + j] = t / d;
    }
    for (int i = 0; i < j; ++i) A[(size_t)i * m + j] = 0.0;
  }
  return A;
}

// inverse of an SPD m x m matrix (row-major) through its Cholesky factor
std::vector<double> spd_inverse(const std::vector<double>& A, int m, const char* what) {
  const std::vector<double> L = small_cholesky(A, m, what);
  std::vector<double> inv((size_t)m * m, 0.0), col((size_t)m);
  for (int c = 0; c < m; ++c) {
    for (int i = 0; i < m; ++i) {  // L y = e_c
      double s = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * m + k] * col[k];
      col[i] = s / L[(size_t)i * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {  // L^T z = y
      double s = col[i];
      for (int k = i + 1; k < m; ++k) s -= L[(size_t)k * m + i] * col[k];
      col[i] = s / L[(size_t)i * m + i];
    }
    for (int i = 0; i < m; ++i) inv[(size_t)i * m + c] = col[i];
  }
  return inv;
}

}  // namespace

struct CholeskySampler {
  int kind = MGMC_FACTOR_SPARSE;
  int64_t n = 0;
  int m = 0, K = 1, nthreads = 1;
  uint32_t k0 = 0, k1 = 0, sample = 0;
  double seconds = 0.0;
  SupernodalFactor L;
  std::vector<double> QV, QWm, QWt;  // low-rank: Q_V, Q_V (W - Id), Q_V (W^T - Id); n x m column-major (factor order)
  std::vector<double> g;             // fixed L^{-1} P f (+ low-rank part), factor order
  bool rhs_fixed = false;

  // y (n x r row-major) += M (Q_V^T y) with M one of QWm / QWt
  void lowrank_apply(const std::vector<double>& M, double* y, int r) const {
    std::vector<double> t((size_t)m * r, 0.0);
    for (int c = 0; c < m; ++c) {
      const double* q = QV.data() + (size_t)c * n;
      double* tc = t.data() + (size_t)c * r;
      for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < r; ++j) tc[j] += q[i] * y[i * r + j];
    }
    for (int c = 0; c < m; ++c) {
      const double* w = M.data() + (size_t)c * n;
      const double* tc = t.data() + (size_t)c * r;
      for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < r; ++j) y[i * r + j] += w[i] * tc[j];
    }
  }

  // y (factor order, n x r) = L^{-1} P f (+ the low-rank part)
  void rhs_part(double* y, int r) const {
    L.solve_L(y, r);
    if (kind == MGMC_FACTOR_LOWRANK && m > 0) lowrank_apply(QWt, y, r);
  }

  void normals(int chain, uint32_t smp, double* out, int r, int c) const {
    for (int64_t v = 0; v < n; v += 4) {
      const u32x4 w = philox4x32_10((uint32_t)(v >> 2), (uint32_t)chain, kCholeskyTag, smp, k0, k1);
      double z[4];
      for (int h = 0; h < 2; ++h) {
        const double u1 = ((double)w.v[2 * h] + 0.5) * (1.0 / 4294967296.0);
        const double u2 = ((double)w.v[2 * h + 1] + 0.5) * (1.0 / 4294967296.0);
        const double rad = std::sqrt(-2.0 * std::log(u1));
        z[2 * h] = rad * std::cos(2.0 * M_PI * u2);
        z[2 * h + 1] = rad * std::sin(2.0 * M_PI * u2);
      }
      for (int e = 0; e < 4 && v + e < n; ++e) out[(v + e) * r + c] = z[e];
    }
  }
};

}  // namespace mgmc

using namespace mgmc;

struct mgmc_cholesky {
  CholeskySampler S;
};

#define CHOL_BEGIN try {
#define CHOL_END                     \
  }                                  \
  catch (const Error& e) {           \
    set_last_error(e.what());        \
    return e.code;                   \
  }                                  \
  catch (const std::exception& e) {  \
    set_last_error(e.what());        \
    return MGMC_ERR_ARG;             \
  }                                  \
  return MGMC_OK;

extern "C" {

int mgmc_cholesky_create(const mgmc_problem* prob, int factorisation, int nchains, uint64_t seed, int nthreads,
                         mgmc_cholesky** out) {
  CHOL_BEGIN
  if (!prob || !out) throw Error(MGMC_ERR_ARG, "null argument");
  *out = nullptr;
  if (factorisation != MGMC_FACTOR_SPARSE && factorisation != MGMC_FACTOR_LOWRANK &&
      factorisation != MGMC_FACTOR_DENSE)
    throw Error(MGMC_ERR_ARG, "factorisation must be sparse, lowrank or dense");
  if (nchains < 1) throw Error(MGMC_ERR_ARG, "nchains must be >= 1");
  const auto t0 = std::chrono::steady_clock::now();
  // the fine prior operator from the library's own assembly (constant,
  // per-vertex or CSR-given stencils)
  HostProblem hp{prob->dim,     {prob->n[0], prob->n[1], prob->n[2]},
                 prob->kappa_sq, prob->A_rowptr, prob->A_col, prob->A_val, 0, nullptr, nullptr, nullptr, nullptr,
                 prob->diffusion, prob->kappa_sq_field};
  const HostHierarchy H = build_hierarchy(hp, 1, 1.0, false);
  const HostLevel& L0 = H.levels[0];
  const Geom& g = L0.g;
  const int64_t n = g.N;
  if (n >= (int64_t(1) << 31)) throw Error(MGMC_ERR_UNSUPPORTED, "lattice too large for the Cholesky samplers");
  const int m = prob->m;
  if (m < 0) throw Error(MGMC_ERR_ARG, "m must be >= 0");
  if (m > 0 && (!prob->B_rowptr || !prob->B_col || !prob->B_val || !prob->sigma))
    throw Error(MGMC_ERR_ARG, "m > 0 needs B and Sigma");
  if (factorisation == MGMC_FACTOR_DENSE && n > 16384)
    throw Error(MGMC_ERR_ARG, "the dense Cholesky sampler is limited to 16384 unknowns");
  std::vector<std::vector<std::pair<int32_t, double>>> rows((size_t)n);
  const int ext[3] = {g.n[0] - 1, g.n[1] - 1, g.dim == 3 ? g.n[2] - 1 : 1};
  for (int64_t r = 0; r < n; ++r) {
    int e[3] = {0, 0, 0};
    g.euclid(r, e);
    const int64_t p = g.padded(r);
    for (int s = 0; s < 27; ++s) {
      const int o[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
      if (g.dim == 2 && o[2] != 0) continue;
      const double a = L0.coef(p, s);
      if (a == 0.0) continue;
      int64_t col = 0, stride = 1;
      bool inside = true;
      for (int d = 0; d < g.dim; ++d) {
        const int b = e[d] + o[d];
        inside = inside && b >= 1 && b <= ext[d];
        col += (int64_t)(b - 1) * stride;
        stride *= ext[d];
      }
      if (inside) rows[(size_t)r].push_back({(int32_t)col, a});
    }
  }
  std::vector<std::vector<std::pair<int32_t, double>>> Bcols((size_t)m);
  for (int64_t r = 0; r < n && m > 0; ++r)
    for (int64_t k = prob->B_rowptr[r]; k < prob->B_rowptr[r + 1]; ++k) {
      const int64_t c = prob->B_col[k];
      if (c < 0 || c >= m) throw Error(MGMC_ERR_ARG, "B column index out of range");
      Bcols[(size_t)c].push_back({(int32_t)r, prob->B_val[k]});
    }
  for (int k = 0; k < m; ++k)
    if (!(prob->sigma[k] > 0.0)) throw Error(MGMC_ERR_ARG, "measurement variances must be positive");
  auto h = std::make_unique<mgmc_cholesky>();
  CholeskySampler& S = h->S;
  S.kind = factorisation;
  S.n = n;
  S.m = m;
  S.K = nchains;
  S.k0 = (uint32_t)seed;
  S.k1 = (uint32_t)(seed >> 32);
  S.nthreads = nthreads > 0 ? nthreads : std::max(1u, std::thread::hardware_concurrency());
  if (factorisation != MGMC_FACTOR_LOWRANK && m > 0) {  // Q = A + B Sigma^{-1} B^T
    int64_t extra = 0;
    for (int k = 0; k < m; ++k) extra += (int64_t)Bcols[k].size() * (int64_t)Bcols[k].size();
    if (extra > kMaxCliqueEntries)
      throw Error(MGMC_ERR_ARG, "B Sigma^{-1} B^T has too many entries for the sparse sampler; use lowrank");
    for (int k = 0; k < m; ++k)
      for (auto& a : Bcols[k])
        for (auto& b : Bcols[k]) rows[(size_t)a.first].push_back({b.first, a.second * b.second / prob->sigma[k]});
  }
  const Csr A = to_csr(rows);
  const bool dense = factorisation == MGMC_FACTOR_DENSE;
  std::vector<std::vector<int32_t>> cliques;
  if (factorisation == MGMC_FACTOR_SPARSE)
    for (int k = 0; k < m; ++k) {
      cliques.emplace_back();
      for (auto& e : Bcols[k]) cliques.back().push_back(e.first);
    }
  std::vector<int32_t> ord = dense ? std::vector<int32_t>() : nested_dissection(g, cliques);
  S.L.analyse(A, ord, dense);
  S.L.factorise(A, S.nthreads);
  if (factorisation == MGMC_FACTOR_LOWRANK && m > 0) {
    // V = L^{-1} P B (n x m), built row-major for the multi-column solve
    std::vector<double> Vr((size_t)n * m, 0.0);
    for (int k = 0; k < m; ++k)
      for (auto& e : Bcols[k]) Vr[(size_t)S.L.iperm[e.first] * m + k] += e.second;
    S.L.solve_L(Vr.data(), m);
    std::vector<double> V((size_t)n * m), R;
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < m; ++k) V[(size_t)k * n + i] = Vr[(size_t)i * m + k];
    std::vector<double>().swap(Vr);
    std::vector<double> C((size_t)m * m, 0.0);  // Sigma + V^T V
    for (int a = 0; a < m; ++a)
      for (int b = 0; b <= a; ++b) {
        double s = 0.0;
        const double* va = V.data() + (size_t)a * n;
        const double* vb = V.data() + (size_t)b * n;
        for (int64_t i = 0; i < n; ++i) s += va[i] * vb[i];
        C[(size_t)a * m + b] = C[(size_t)b * m + a] = s + (a == b ? prob->sigma[a] : 0.0);
      }
    householder_qr(n, m, std::move(V), S.QV, R);
    const std::vector<double> Ci = spd_inverse(C, m, "Sigma + V^T V");
    std::vector<double> M((size_t)m * m, 0.0), RC((size_t)m * m, 0.0);  // Id - R C^{-1} R^T
    for (int a = 0; a < m; ++a)
      for (int q = 0; q < m; ++q) {
        double s = 0.0;
        for (int p = 0; p < m; ++p) s += R[(size_t)a * m + p] * Ci[(size_t)p * m + q];
        RC[(size_t)a * m + q] = s;
      }
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        double s = 0.0;
        for (int q = 0; q < m; ++q) s += RC[(size_t)a * m + q] * R[(size_t)b * m + q];
        M[(size_t)a * m + b] = (a == b ? 1.0 : 0.0) - s;
      }
    for (int a = 0; a < m; ++a)  // symmetrise the rounding
      for (int b = 0; b < a; ++b) M[(size_t)a * m + b] = M[(size_t)b * m + a] = 0.5 * (M[(size_t)a * m + b] + M[(size_t)b * m + a]);
    const std::vector<double> W = small_cholesky(M, m, "Id - Lambda");
    S.QWm.assign((size_t)n * m, 0.0);
    S.QWt.assign((size_t)n * m, 0.0);
    for (int c = 0; c < m; ++c)
      for (int p = 0; p < m; ++p) {
        const double wl = W[(size_t)p * m + c] - (p == c ? 1.0 : 0.0);  // (W - Id)(p, c)
        const double wu = W[(size_t)c * m + p] - (p == c ? 1.0 : 0.0);  // (W^T - Id)(p, c)
        const double* q = S.QV.data() + (size_t)p * n;
        double* a = S.QWm.data() + (size_t)c * n;
        double* b = S.QWt.data() + (size_t)c * n;
        for (int64_t i = 0; i < n; ++i) {
          a[i] += q[i] * wl;
          b[i] += q[i] * wu;
        }
      }
  }
  S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = h.release();
  CHOL_END
}

int mgmc_cholesky_destroy(mgmc_cholesky* c) {
  delete c;
  return MGMC_OK;
}

int mgmc_cholesky_info(const mgmc_cholesky* c, int64_t* n, int64_t* nnz_L, int64_t* nsupernodes, double* seconds) {
  CHOL_BEGIN
  if (!c) throw Error(MGMC_ERR_ARG, "null handle");
  if (n) *n = c->S.n;
  if (nnz_L) *nnz_L = c->S.L.nnz;
  if (nsupernodes) *nsupernodes = c->S.L.ns;
  if (seconds) *seconds = c->S.seconds;
  CHOL_END
}

int mgmc_cholesky_set_nchains(mgmc_cholesky* c, int nchains) {
  CHOL_BEGIN
  if (!c) throw Error(MGMC_ERR_ARG, "null handle");
  if (nchains < 1) throw Error(MGMC_ERR_ARG, "nchains must be >= 1");
  c->S.K = nchains;
  CHOL_END
}

int mgmc_cholesky_fix_rhs(mgmc_cholesky* c, const double* f) {
  CHOL_BEGIN
  if (!c || !f) throw Error(MGMC_ERR_ARG, "null argument");
  CholeskySampler& S = c->S;
  std::vector<double> y((size_t)S.n);
  for (int64_t i = 0; i < S.n; ++i) y[i] = f[S.L.perm[i]];
  S.rhs_part(y.data(), 1);
  S.g.swap(y);
  S.rhs_fixed = true;
  CHOL_END
}

int mgmc_cholesky_unfix_rhs(mgmc_cholesky* c) {
  CHOL_BEGIN
  if (!c) throw Error(MGMC_ERR_ARG, "null handle");
  c->S.rhs_fixed = false;
  CHOL_END
}

int mgmc_cholesky_sample(mgmc_cholesky* c, const double* f, int f_per_chain, const double* xi, double* x) {
  CHOL_BEGIN
  if (!c || !x) throw Error(MGMC_ERR_ARG, "null argument");
  CholeskySampler& S = c->S;
  if (!S.rhs_fixed && !f) throw Error(MGMC_ERR_ARG, "f is needed unless the right-hand side is fixed");
  const uint32_t smp = S.sample++;
  const int64_t n = S.n;
  const int64_t nblk = (S.K + kRhsBlock - 1) / kRhsBlock;
  parallel_for(S.nthreads, nblk, [&](int64_t blk) {
    const int c0 = (int)(blk * kRhsBlock), r = std::min(kRhsBlock, S.K - c0);
    std::vector<double> y((size_t)n * r), z((size_t)n * r);
    if (S.rhs_fixed) {
      for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < r; ++j) y[i * r + j] = S.g[i];
    } else {
      for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < r; ++j) y[i * r + j] = f[(f_per_chain ? (size_t)(c0 + j) * n : 0) + S.L.perm[i]];
      S.rhs_part(y.data(), r);
    }
    for (int j = 0; j < r; ++j) {
      if (xi) {
        const double* xc = xi + (size_t)(c0 + j) * n;
        for (int64_t i = 0; i < n; ++i) z[i * r + j] = xc[i];
      } else {
        S.normals(c0 + j, smp, z.data(), r, j);
      }
    }
    for (size_t e = 0; e < z.size(); ++e) z[e] += y[e];
    if (S.kind == MGMC_FACTOR_LOWRANK && S.m > 0) S.lowrank_apply(S.QWm, z.data(), r);
    S.L.solve_LT(z.data(), r);
    for (int j = 0; j < r; ++j) {
      double* xc = x + (size_t)(c0 + j) * n;
      for (int64_t i = 0; i < n; ++i) xc[S.L.perm[i]] = z[i * r + j];
    }
  });
  CHOL_END
}

}  // extern "C"
