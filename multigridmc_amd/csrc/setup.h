// Host-side (FP64) construction of the multigrid hierarchy the kernels run on.
//
// Device data layout. Level l of a lattice with n_d cells per direction keeps
// every field as a zero-padded box that includes the Dirichlet boundary
// vertices: index (k * py + j) * px + i with 0 <= i <= n_x etc., pitch
// px = round_up(n_x, 8) (aligned 8-vertex groups for the noise and the
// vector loads), py = n_y + 1, pz = n_z + 1 (1 in 2D). When n_x is a
// multiple of 8 the right boundary vertex i = n_x of row j shares its slot
// with the left boundary vertex i = 0 of row j + 1 (both are zero), so a row
// holds no padding at all. Boundary and padding entries stay zero, so a
// constant stencil applied everywhere equals the reference's operator with
// eliminated boundary rows; kernels write boundary slots only with zero.
// K chains are stored back to back with a stride rounded up to 64 elements;
// the slot after the last vertex of a chain is the next chain's first
// (boundary) slot, inter-chain padding or the field's trailing zero guard.
#ifndef MGMC_SETUP_H
#define MGMC_SETUP_H

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace mgmc {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Geom {
  int dim = 3;
  int n[3] = {1, 1, 1};
  int px = 0, py = 0, pz = 0;
  int64_t vol = 0;     // px * py * pz
  int64_t cstride = 0; // per-chain stride (elements)
  int64_t N = 0;       // interior unknowns

  static Geom make(int dim, const int* n);
  Geom coarse() const;
  // reference vertex index -> padded index
  int64_t padded(int64_t ell) const;
  void euclid(int64_t ell, int* idx) const;
};

// Sparse rows: x[row[r]] += sum_{e in [ptr[r], ptr[r+1])} val[e] * u[col[e]]
struct SparseRows {
  std::vector<int64_t> row;
  std::vector<int32_t> ptr{0};
  std::vector<int32_t> col;
  std::vector<double> val;
  int nrows() const { return (int)row.size(); }
};

// Offsets of the stencil planes of a variable star level, in stencil-index
// order: -z, -y, -x, centre, +x, +y, +z.
constexpr int kStarPlanes = 7;
constexpr int kStarIndex[kStarPlanes] = {4, 10, 12, 13, 14, 16, 22};

struct HostLevel {
  Geom g;
  bool star = true;      // 5/7-point stencil -> red-black; else 2^d colours
  int ncolour = 2;
  double st[27] = {0};   // stencil, index (dz+1)*9 + (dy+1)*3 + (dx+1)
  double diag = 0;
  // Variable coefficients: a stencil per vertex, stored as planes of vol
  // values in the padded layout (vst[plane * vol + p]); 7 planes
  // (kStarIndex) on star levels, 27 on box levels; zero at boundary slots.
  // st / diag are unused then.
  bool variable = false;
  std::vector<double> vst;
  int nplanes() const { return variable ? (star ? kStarPlanes : 27) : 0; }
  // coefficient e (stencil index) of the vertex at padded index p
  double coef(int64_t p, int e) const {
    if (!variable) return st[e];
    if (!star) return vst[(size_t)e * g.vol + p];
    for (int q = 0; q < kStarPlanes; ++q)
      if (kStarIndex[q] == e) return vst[(size_t)q * g.vol + p];
    return 0.0;
  }
  int m = 0;
  // columns of B_l as (padded index, value), used for z = B^T x
  std::vector<int32_t> bptr{0};
  std::vector<int64_t> bidx;
  std::vector<double> bval;
  SparseRows Brows;      // B_l by rows (coarse-side term of the residual)
  SparseRows W[2];       // W = (L_colour + D/omega)^{-1} B, [0] backward, [1] forward
  std::vector<double> Sinv[2];  // (Sigma + B^T W)^{-1}, m x m row-major
  std::vector<double> G[2];     // B^T W, m x m row-major
};

struct HostCoarse {
  int64_t Nc = 0;
  std::vector<int64_t> pidx;    // padded index of coarse vertex (reference order)
  std::vector<double> LinvT;    // L^{-T}, Nc x Nc row-major
  std::vector<double> Qinv;     // Q_c^{-1} = L^{-T} L^{-1}
};

struct HostProblem {
  int dim;
  int n[3];
  double kappa_sq;
  const int64_t* A_rowptr;
  const int64_t* A_col;
  const double* A_val;
  int m;
  const int64_t* B_rowptr;
  const int64_t* B_col;
  const double* B_val;
  const double* sigma;
  const double* diffusion = nullptr;       // cell-wise a (prod n_d values, x fastest) or NULL
  const double* kappa_sq_field = nullptr;  // vertex-wise kappa^2 (reference order) or NULL
};

struct HostHierarchy {
  std::vector<HostLevel> levels;
  HostCoarse coarse;
  std::vector<double> sigma;
  double omega = 1.0;
};

// Ball-average measurement vector (reference vertex indices ascending, values)
// and the Q1 FEM stencil; operators.cc.
// profile 0: top-hat (the reference's), 1: Gaussian weight of the given width
// (standard deviation in units of the radius), both of integral one on the ball.
void measurement_vector(int dim, const int* n, const double* x0, double radius, std::vector<int64_t>& idx,
                        std::vector<double>& val, int profile = 0, double width = 0.5);
double gaussian_ball_integral(double width, int dim);
void fem_stencil(int dim, const int* n, double kappa_sq, double* st);

// Largest coarsest level (unknowns) for the dense Cholesky factor; the
// coarse kernel stages 2 Nc doubles in LDS, which bounds it by kMaxCoarseKernel.
constexpr int64_t kMaxCoarse = 8192;
constexpr int64_t kMaxCoarseKernel = 2048;
// coarse_cholesky: the coarsest level is sampled with the Cholesky sampler
// (the factor is then required); otherwise it is built when it fits.
HostHierarchy build_hierarchy(const HostProblem& p, int nlevel, double omega, bool coarse_cholesky = true);

// Message returned by mgmc_last_error() on this thread (mgmc_lib.hip).
void set_last_error(const std::string& message);

// (Sigma + B^T T^{-1} B)^{-1} of a level for the lexicographic visiting order
// ([0] backward, [1] forward; m x m row-major), T = D/omega + the couplings to
// the vertices visited before.
void lex_capacitance(const HostLevel& L, double omega, const std::vector<double>& sigma, std::vector<double> out[2]);

// Galerkin coarse stencil of a constant fine stencil with linear interpolation.
void galerkin_stencil(int dim, const double* fine, double* coarse);

}  // namespace mgmc

#endif
