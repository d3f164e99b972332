// Kernels of the MGMC cycle for gfx950 (64-wide wavefronts).
//
// Sweeps are multicolour SOR/Gibbs updates: red-black for star stencils
// (5/7-point), 2^d colours for box stencils (9/27-point, every Galerkin
// level). One thread owns an aligned group of eight vertices of one lattice
// row and updates the four of them that carry the current colour, so that a
// single Philox call supplies its four normals and the row segment moves as
// whole 32-byte (FP32) / 64-byte (FP64) vectors.
#include "kernels.h"

namespace mgmc {

template <typename T>
__device__ __forceinline__ void load8(const T* __restrict__ p, T* v) {
  if constexpr (sizeof(T) == 4) {
    const float4* q = reinterpret_cast<const float4*>(p);
    float4 a = q[0], b = q[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double2 a = q[i];
      v[2 * i] = a.x;
      v[2 * i + 1] = a.y;
    }
  }
}

template <typename T>
__device__ __forceinline__ void store8(T* __restrict__ p, const T* v) {
  if constexpr (sizeof(T) == 4) {
    float4* q = reinterpret_cast<float4*>(p);
    q[0] = make_float4(v[0], v[1], v[2], v[3]);
    q[1] = make_float4(v[4], v[5], v[6], v[7]);
  } else {
    double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = make_double2(v[2 * i], v[2 * i + 1]);
  }
}

// LDS copy of an x-row of px = 8 ngx values, as 8 / C chunk planes of C =
// 16 / sizeof(T) values per group: chunk h of group g sits at h C ngx + C g,
// so the 16-byte accesses of the lanes of a row are consecutive (a plain
// 8-value stride leaves half of the LDS banks idle and conflicts 2-way).
template <typename T>
struct LdsRow {
  static constexpr int C = 16 / (int)sizeof(T), NCH = 8 / C;
  __device__ static __forceinline__ void store(T* row, int gx, int ngx, const T* v) {
#pragma unroll
    for (int h = 0; h < NCH; ++h) {
      T* q = row + h * C * ngx + C * gx;
      if constexpr (sizeof(T) == 4)
        *reinterpret_cast<float4*>(q) = make_float4(v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
      else
        *reinterpret_cast<double2*>(q) = make_double2(v[2 * h], v[2 * h + 1]);
    }
  }
  __device__ static __forceinline__ void load(const T* row, int gx, int ngx, T* v) {
#pragma unroll
    for (int h = 0; h < NCH; ++h) {
      const T* q = row + h * C * ngx + C * gx;
      if constexpr (sizeof(T) == 4) {
        const float4 a = *reinterpret_cast<const float4*>(q);
        v[4 * h] = a.x; v[4 * h + 1] = a.y; v[4 * h + 2] = a.z; v[4 * h + 3] = a.w;
      } else {
        const double2 a = *reinterpret_cast<const double2*>(q);
        v[2 * h] = a.x; v[2 * h + 1] = a.y;
      }
    }
  }
  // x(8 gx - 1) and x(8 gx + 8): the last value of group gx - 1, the first of gx + 1
  __device__ static __forceinline__ T left(const T* row, int gx, int ngx) {
    return row[(NCH - 1) * C * ngx + C * (gx - 1) + C - 1];
  }
  __device__ static __forceinline__ T right(const T* row, int gx, int ngx) { return row[C * (gx + 1)]; }
};

// Star-stencil update of the four vertices of parity IPAR in an 8-vertex
// group: xc holds the group, `left`/`right` its outer neighbours, ym/yp/zm/zp
// the same group in the neighbouring rows/planes, fb the right-hand side.
// IPAR is a template parameter so every index below is a compile-time
// constant and the groups stay in registers.
template <int IPAR, typename T, int NOISE>
__device__ __forceinline__ void star_update4(T* xc, T left, T right, const T* ym, const T* yp, const T* zm,
                                             const T* zp, const T* fb, const float* xi, const T* extg, int i0, int n0,
                                             T cx, T cy, T cz, T D, T om_d, T ns) {
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int q = 2 * l + IPAR;
    const T xl = q > 0 ? xc[q - 1] : left;
    const T xr = q < 7 ? xc[q + 1] : right;
    T acc = cy * ym[q] + cx * xl + D * xc[q] + cx * xr + cy * yp[q];
    if (zm) acc = cz * zm[q] + acc + cz * zp[q];
    const int i = i0 + q;
    if (i >= 1 && i <= n0 - 1) {
      T bb = fb[q];
      if (NOISE == NOISE_PHILOX) bb += ns * (T)xi[l];
      if (NOISE == NOISE_EXTERNAL) bb += ns * extg[q];
      xc[q] += om_d * (bb - acc);
    }
  }
}

// Box-stencil (9/27-point) update of the four vertices of parity IPAR in an
// 8-vertex group; xc[q + 1] = x(i0 + q), xc[0] and xc[9] the outer
// neighbours, xr the group's address (neighbour rows are loaded from it).
template <int IPAR, typename T, int DIM, int NOISE>
__device__ __forceinline__ void box_update4(T* xc, const T* xr, const DevGeom& g, const Stencil& st, const T* fb,
                                            const float* xi, const T* extg, int i0) {
  T acc[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) acc[l] = (T)0;
  const long long pl = (long long)g.px * g.py;
#pragma unroll
  for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz) {
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      T w[10];
      if (dz == 0 && dy == 0) {
#pragma unroll
        for (int q = 0; q < 10; ++q) w[q] = xc[q];
      } else {
        const T* p = xr + dz * pl + dy * g.px;
        load8<T>(p, w + 1);
        w[0] = p[-1];
        w[9] = p[8];
      }
      const T a0 = st.c[(dz + 1) * 9 + (dy + 1) * 3 + 0];
      const T a1 = st.c[(dz + 1) * 9 + (dy + 1) * 3 + 1];
      const T a2 = st.c[(dz + 1) * 9 + (dy + 1) * 3 + 2];
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int q = 2 * l + IPAR;
        acc[l] += a0 * w[q] + a1 * w[q + 1] + a2 * w[q + 2];
      }
    }
  }
  const T om_d = st.inv_diag_omega;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int q = 2 * l + IPAR;
    const int i = i0 + q;
    if (i < 1 || i > g.n0 - 1) continue;
    T b = fb[q];
    if (NOISE == NOISE_PHILOX) b += st.noise_scale * (T)xi[l];
    if (NOISE == NOISE_EXTERNAL) b += st.noise_scale * extg[q];
    xc[q + 1] += om_d * (b - acc[l]);
  }
}

// Box-stencil update of the four parity-IPAR vertices of a group whose own
// plane is in LDS, xc its values with margins. CHUNK: `row`, `zm`, `zp` are
// the starts of the group's row in LDS planes k, k-1, k+1 stored as LdsRow
// chunks; otherwise they point at the group itself in plain rows (LDS or
// global memory). Planes are visited one at a time (one plane's rows live:
// 62 VGPRs in FP32, no spills in FP64); the centre row is read back from
// LDS, where it still holds the values of this phase. Outer margins of the
// first / last group are never used.
template <int IPAR, typename T, int NOISE, bool CHUNK>
__device__ __forceinline__ void box_plane_update4(T* xc, const T* row, const T* zm, const T* zp, const DevGeom& g,
                                                  const StencilT<T>& st, const T* __restrict__ coef, const T* fb,
                                                  const float* xi, const T* extg, int i0, int gx, int ngx) {
  T acc[4];
#pragma unroll
  for (int l = 0; l < 4; ++l) acc[l] = (T)0;
#pragma unroll 1
  for (int dz = -1; dz <= 1; ++dz) {
    const T* pl = dz < 0 ? zm : dz > 0 ? zp : row;
    const T* cz = coef + (dz + 1) * 9;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      T w[10];
      const T* p = pl + dy * g.px;
      if constexpr (CHUNK) {
        LdsRow<T>::load(p, gx, ngx, w + 1);
        w[0] = gx > 0 ? LdsRow<T>::left(p, gx, ngx) : (T)0;
        w[9] = gx < ngx - 1 ? LdsRow<T>::right(p, gx, ngx) : (T)0;
      } else {
        load8<T>(p, w + 1);
        w[0] = gx > 0 ? p[-1] : (T)0;
        w[9] = gx < ngx - 1 ? p[8] : (T)0;
      }
      const T a0 = cz[(dy + 1) * 3 + 0];
      const T a1 = cz[(dy + 1) * 3 + 1];
      const T a2 = cz[(dy + 1) * 3 + 2];
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int q = 2 * l + IPAR;
        acc[l] += a0 * w[q] + a1 * w[q + 1] + a2 * w[q + 2];
      }
    }
  }
  const T om_d = st.inv_diag_omega;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int q = 2 * l + IPAR;
    const int i = i0 + q;
    if (i < 1 || i > g.n0 - 1) continue;
    T b = fb[q];
    if (NOISE == NOISE_PHILOX) b += st.noise_scale * (T)xi[l];
    if (NOISE == NOISE_EXTERNAL) b += st.noise_scale * extg[q];
    xc[q + 1] += om_d * (b - acc[l]);
  }
}

// x += alpha P xc for the fine 8-vertex group (8 gx .. 8 gx + 7, row j) of a
// thread that visits the fine planes of its row in increasing order (the
// gather form of k_prolongate, fused into the loads of a sweep). Fine
// boundary vertices receive only coarse boundary values, which are zero.
// The thread holds the y-interpolated
// coarse values s(K) and s(K+1) of the two coarse planes around the current
// fine plane, so every coarse plane is read once per thread (two aligned
// 4-vertex loads and one single load per coarse row).
template <typename T>
struct ProlongStream {
  const T* cb;  // coarse chain base, nullptr: no correction
  DevGeom gc;
  int gx, j0, j1, kl;
  T alpha;
  T sl[5], sh[5];

  // s = the mean of coarse rows j0 and j1 of plane K (one row when j0 = j1:
  // (v + v) / 2 = v exactly); both rows' loads are issued together
  __device__ __forceinline__ void row_values(int K, T* s) const {
#pragma unroll
    for (int e = 0; e < 5; ++e) s[e] = (T)0;
    if (K > gc.n2) return;  // beyond the last coarse plane: boundary, zero
    T v[2][5];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const T* p = cb + ((long long)K * gc.py + (b ? j1 : j0)) * gc.px + 4 * gx;
      if constexpr (sizeof(T) == 4) {
        const float4 a = *reinterpret_cast<const float4*>(p);
        v[b][0] = a.x; v[b][1] = a.y; v[b][2] = a.z; v[b][3] = a.w;
      } else {
        const double2 a = reinterpret_cast<const double2*>(p)[0], c = reinterpret_cast<const double2*>(p)[1];
        v[b][0] = a.x; v[b][1] = a.y; v[b][2] = c.x; v[b][3] = c.y;
      }
      v[b][4] = 4 * gx + 4 < gc.px ? p[4] : (T)0;
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) s[e] = (T)0.5 * (v[0][e] + v[1][e]);
  }
  // first fine plane this thread will visit
  __device__ __forceinline__ void init(const T* coarse, const DevGeom& g_c, int gx_, int j, int p0, T a) {
    cb = coarse;
    gc = g_c;
    gx = gx_;
    j0 = j >> 1;
    j1 = (j + 1) >> 1;
    alpha = a;
    kl = p0 >> 1;
    if (!cb) return;
    row_values(kl, sl);
    row_values(kl + 1, sh);
  }
  // v += alpha (P xc) on fine plane p (p never below the previous call's)
  __device__ __forceinline__ void add(T* v, int p) {
    advance(p);
    apply(v, p);
  }
  // the two halves of add: advance(p) requests the coarse rows fine plane p
  // needs (when the thread requests plane p of x), apply(v, p) adds the
  // correction (when plane p is used), so that both loads are in flight
  // meanwhile; between them no other plane may be advanced to
  __device__ __forceinline__ void advance(int p) {
    if (!cb) return;
    if ((p >> 1) != kl) {
      ++kl;
#pragma unroll
      for (int e = 0; e < 5; ++e) sl[e] = sh[e];
      row_values(kl + 1, sh);
    }
  }
  __device__ __forceinline__ void apply(T* v, int p) const {
    if (!cb) return;
    T t[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) t[e] = (p & 1) ? (T)0.5 * (sl[e] + sh[e]) : sl[e];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] += alpha * ((q & 1) ? (T)0.5 * (t[q >> 1] + t[(q >> 1) + 1]) : t[q >> 1]);
  }
};


// ---------------------------------------------------------------------------
// One colour phase of an SOR sweep (SORSmoother::apply_sparse,
// src/smoother/sor_smoother.cc:56-77, restricted to one colour), with the
// Gibbs right-hand side c = f + sqrt(D(2-omega)/omega) xi of
// SORSampler::apply (src/sampler/sor_sampler.cc:42-48) when NOISE != 0.
// STAR: colour = (i+j+k)&1; box: colour = (i&1) | (j&1)<<1 | (k&1)<<2.
// ---------------------------------------------------------------------------
template <typename T, int DIM, bool STAR, int NOISE>
__device__ __forceinline__ void b_sweep(const Blk bk, T* __restrict__ x, const T* __restrict__ f,
                                               const T* __restrict__ ext, DevGeom g, Stencil st, int colour,
                                               int nrows, int K, NoiseKey nk) {
  const int ngx = g.px >> 3;
  long long t = (long long)bk.x * blockDim.x + threadIdx.x;
  int gx = (int)(t % ngx);
  long long row = t / ngx;
  // groups holding no interior vertex (the x padding) change nothing
  if (row >= (long long)nrows * K || 8 * gx > g.n0 - 1) return;
  int chain = (int)(row / nrows);
  int rr = (int)(row % nrows);
  int j, k, ipar;
  if (STAR) {
    j = rr % (g.n1 - 1) + 1;
    k = DIM == 3 ? rr / (g.n1 - 1) + 1 : 0;
    ipar = (colour + j + k) & 1;
  } else {
    int jpar = (colour >> 1) & 1, kpar = (colour >> 2) & 1;
    int nj = jpar ? g.n1 / 2 : (g.n1 - 1) / 2;  // odd rows 1,3,.. / even rows 2,4,.. below n1
    int jj = rr % nj;
    j = 2 * jj + (jpar ? 1 : 2);
    if (DIM == 3) {
      int kk = rr / nj;
      k = 2 * kk + (kpar ? 1 : 2);
    } else {
      k = 0;
    }
    ipar = colour & 1;
  }
  const long long rowoff = ((long long)k * g.py + j) * g.px;
  const long long base = chain * g.cs + rowoff + 8 * gx;
  T* xr = x + base;

  // centre row with one-vertex margins: xc[q + 1] = x(8 gx + q)
  T xc[10];
  load8<T>(xr, xc + 1);
  xc[0] = xr[-1];
  xc[9] = xr[8];
  T fb[8];
  load8<T>(f + chain * g.fcs + rowoff + 8 * gx, fb);

  float xi[4] = {0.f, 0.f, 0.f, 0.f};
  if (NOISE == NOISE_PHILOX) {
    uint32_t c0 = (uint32_t)((((long long)k * g.py + j) * ngx + gx) << 1) | (uint32_t)ipar;
    normals4(c0, (uint32_t)chain, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
  }

  if (STAR) {
    T ym[8], yp[8], zm[8], zp[8];
    load8<T>(xr - g.px, ym);
    load8<T>(xr + g.px, yp);
    if (DIM == 3) {
      const long long pl = (long long)g.px * g.py;
      load8<T>(xr - pl, zm);
      load8<T>(xr + pl, zp);
    }
    const T cx = (T)st.c[12], cy = (T)st.c[10], cz = (T)st.c[4];
    const T* eg = NOISE == NOISE_EXTERNAL ? ext + base : nullptr;
    const T* zmp = DIM == 3 ? zm : nullptr;
    if (ipar == 0)
      star_update4<0, T, NOISE>(xc + 1, xc[0], xc[9], ym, yp, zmp, zp, fb, xi, eg, 8 * gx, g.n0, cx, cy, cz,
                                (T)st.diag, (T)st.inv_diag_omega, (T)st.noise_scale);
    else
      star_update4<1, T, NOISE>(xc + 1, xc[0], xc[9], ym, yp, zmp, zp, fb, xi, eg, 8 * gx, g.n0, cx, cy, cz,
                                (T)st.diag, (T)st.inv_diag_omega, (T)st.noise_scale);
    store8<T>(xr, xc + 1);
    return;
  }
  const T* eg = NOISE == NOISE_EXTERNAL ? ext + base : nullptr;
  if (ipar == 0)
    box_update4<0, T, DIM, NOISE>(xc, xr, g, st, fb, xi, eg, 8 * gx);
  else
    box_update4<1, T, DIM, NOISE>(xc, xr, g, st, fb, xi, eg, 8 * gx);
  store8<T>(xr, xc + 1);
}
template <typename T, int DIM, bool STAR, int NOISE>
__global__ void __launch_bounds__(256) k_sweep(T* __restrict__ x, const T* __restrict__ f,
                                               const T* __restrict__ ext, DevGeom g, Stencil st, int colour,
                                               int nrows, int K, NoiseKey nk) {
  b_sweep<T, DIM, STAR, NOISE>(Blk::hw(), x, f, ext, g, st, colour, nrows, K, nk);
}
template <typename T, int DIM, bool STAR, int NOISE>
constexpr Batchable<b_sweep<T, DIM, STAR, NOISE>, 256, false, k_sweep<T, DIM, STAR, NOISE>> kb_sweep{};

// ---------------------------------------------------------------------------
// Variable coefficients: every vertex has its own stencil, stored as planes
// of the level's padded box (vc[plane * vol + p]): 7 planes on star levels
// (-z, -y, -x, centre, +x, +y, +z), 27 on box levels (stencil index order).
// The coefficients are shared by all chains.
// ---------------------------------------------------------------------------
template <typename T, int DIM, bool STAR>
__device__ __forceinline__ T apply_var(const T* __restrict__ xp, const T* __restrict__ vc, long long vol, long long p,
                                       const DevGeom& g) {
  const long long pl = (long long)g.px * g.py;
  T acc = (T)0;
  if constexpr (STAR) {
    if (DIM == 3) acc += vc[p] * xp[-pl];
    acc += vc[vol + p] * xp[-g.px];
    acc += vc[2 * vol + p] * xp[-1];
    acc += vc[3 * vol + p] * xp[0];
    acc += vc[4 * vol + p] * xp[1];
    acc += vc[5 * vol + p] * xp[g.px];
    if (DIM == 3) acc += vc[6 * vol + p] * xp[pl];
  } else {
#pragma unroll
    for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int e = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
          acc += vc[e * vol + p] * xp[dz * pl + dy * g.px + dx];
        }
  }
  return acc;
}

// Chains per thread of the variable-coefficient kernels: the vertex's
// stencil is loaded once and applied to VAR_CPT chains.
constexpr int VAR_CPT = 8;

// One colour phase of an SOR sweep with per-vertex stencils (the update of
// SORSmoother::apply_sparse, src/smoother/sor_smoother.cc:56-77, with the
// Gibbs right-hand side of SORSampler::apply, src/sampler/sor_sampler.cc:42-48:
// c = f + sqrt(D (2 - omega) / omega) xi with the vertex's own D). A thread
// updates one vertex of the colour in VAR_CPT consecutive chains (grid:
// vertices x chain groups); the Philox normal of a vertex is the one the
// constant-stencil kernels draw for it.
template <typename T, int DIM, bool STAR, int NOISE>
__global__ void __launch_bounds__(256) k_sweep_var(T* __restrict__ x, const T* __restrict__ f,
                                                   const T* __restrict__ ext, DevGeom g, const T* __restrict__ vc,
                                                   long long vol, double omega, int colour, int K, NoiseKey nk) {
  const int nj = g.n1 - 1, nrow = nj * (DIM == 3 ? g.n2 - 1 : 1);
  const int half = g.n0 / 2 + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nrow * half) return;
  const int i2 = (int)(t % half);
  const int row = (int)(t / half);
  const int c0 = blockIdx.y * VAR_CPT;
  const int j = row % nj + 1;
  const int k = DIM == 3 ? row / nj + 1 : 0;
  int ipar;
  if (STAR) {
    ipar = (colour + j + k) & 1;
  } else {
    if ((j & 1) != ((colour >> 1) & 1)) return;
    if (DIM == 3 && (k & 1) != ((colour >> 2) & 1)) return;
    ipar = colour & 1;
  }
  const int i = 2 * i2 + ipar;
  if (i < 1 || i > g.n0 - 1) return;
  const long long p = ((long long)k * g.py + j) * g.px + i;
  // the vertex's stencil, once for all chains of the thread
  const int np = STAR ? 7 : 27;
  T cf[27];
#pragma unroll
  for (int e = 0; e < 27; ++e)
    if (e < np && (DIM == 3 || !STAR || (e != 0 && e != 6)) && (DIM == 3 || STAR || (e >= 9 && e < 18)))
      cf[e] = vc[e * vol + p];
  const T D = STAR ? cf[3] : cf[13];
  const T ns = NOISE != NOISE_NONE ? (T)__builtin_sqrt((double)D * (2.0 - omega) / omega) : (T)0;
  const long long pl = (long long)g.px * g.py;
  const int ngx = g.px >> 3;
  const uint32_t slot = (uint32_t)((((long long)k * g.py + j) * ngx + (i >> 3)) << 1) | (uint32_t)(i & 1);
  for (int q = 0; q < VAR_CPT; ++q) {
    const int chain = c0 + q;
    if (chain >= K) break;
    T* xp = x + chain * g.cs + p;
    T acc = (T)0;
    if constexpr (STAR) {
      if (DIM == 3) acc += cf[0] * xp[-pl];
      acc += cf[1] * xp[-g.px];
      acc += cf[2] * xp[-1];
      acc += cf[3] * xp[0];
      acc += cf[4] * xp[1];
      acc += cf[5] * xp[g.px];
      if (DIM == 3) acc += cf[6] * xp[pl];
    } else {
#pragma unroll
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx)
            acc += cf[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] * xp[dz * pl + dy * g.px + dx];
    }
    T c = f[chain * g.fcs + p];
    if (NOISE == NOISE_PHILOX) {
      float xi[4];
      normals4(slot, (uint32_t)chain, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
      c += ns * (T)xi[(i & 7) >> 1];
    }
    if (NOISE == NOISE_EXTERNAL) c += ns * ext[chain * g.cs + p];
    *xp += (T)omega * (c - acc) / D;
  }
}

// r = f - A x (f may be NULL: r = -A x) with per-vertex stencils, one thread
// per interior vertex of every chain.
template <typename T, int DIM, bool STAR>
__global__ void __launch_bounds__(256) k_residual_var(const T* __restrict__ x, const T* __restrict__ f,
                                                      T* __restrict__ r, DevGeom g, const T* __restrict__ vc,
                                                      long long vol, int K) {
  const int nj = g.n1 - 1, nrow = nj * (DIM == 3 ? g.n2 - 1 : 1);
  const int ni = g.n0 - 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nrow * ni) return;
  const int i = (int)(t % ni) + 1;
  const int row = (int)(t / ni);
  const int j = row % nj + 1;
  const int k = DIM == 3 ? row / nj + 1 : 0;
  const long long p = ((long long)k * g.py + j) * g.px + i;
  const int c0 = blockIdx.y * VAR_CPT;
  for (int q = 0; q < VAR_CPT && c0 + q < K; ++q) {
    const int chain = c0 + q;
    const T acc = apply_var<T, DIM, STAR>(x + chain * g.cs + p, vc, vol, p, g);
    r[chain * g.cs + p] = (f ? f[chain * g.fcs + p] : (T)0) - acc;
  }
}

// ---------------------------------------------------------------------------
// One wavefront of a lexicographic SOR sweep: the reference's own visiting
// order (SORSmoother::apply_sparse, src/smoother/sor_smoother.cc:56-77,
// vertices in increasing index order forward, decreasing backward) run in
// parallel. With hyperplanes s = i + ay j + az k, where (ay, az) = (1, 1) for
// star stencils and (2, 4) for box stencils, every neighbour that precedes a
// vertex in index order lies on a lower hyperplane and every later one on a
// higher hyperplane, so updating hyperplane s after s - 1 (forward; s + 1
// backward) sees exactly the old and new values the sequential sweep sees.
// The arithmetic follows the reference's: res = sum of the row's entries times
// x in increasing column order (zero stencil entries and boundary slots add
// exact zeros), x += omega (c - res) / D, c = f + sqrt(D (2 - omega) / omega) xi.
// A thread updates one vertex; (j, k, chain) come from the thread index, i
// from s. The Philox normal of a vertex is the one the multicolour kernels
// draw for it (lane (i & 7) >> 1 of its group's call).
// ---------------------------------------------------------------------------
template <typename T, int DIM, int NOISE>
__global__ void __launch_bounds__(256) k_sweep_lex(T* __restrict__ x, const T* __restrict__ f,
                                                   const T* __restrict__ ext, DevGeom g, Stencil st, int s, int ay,
                                                   int az, int K, NoiseKey nk, const T* __restrict__ vc,
                                                   long long vol, int nplanes) {
  const int nj = g.n1 - 1, nk3 = DIM == 3 ? g.n2 - 1 : 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nj * nk3 * K) return;
  const int j = (int)(t % nj) + 1;
  const long long r = t / nj;
  const int k = DIM == 3 ? (int)(r % nk3) + 1 : 0;
  const int chain = (int)(r / nk3);
  const int i = s - ay * j - az * k;
  if (i < 1 || i > g.n0 - 1) return;
  const long long off = ((long long)k * g.py + j) * g.px + i;
  T* xp = x + chain * g.cs + off;
  const long long pl = (long long)g.px * g.py;
  T res = (T)0;
  double diag = st.diag;
  if (vc) {  // per-vertex stencils: 7 planes (star) or 27
    if (nplanes == 7) {
      if (DIM == 3) res += vc[off] * xp[-pl];
      res += vc[vol + off] * xp[-g.px];
      res += vc[2 * vol + off] * xp[-1];
      res += vc[3 * vol + off] * xp[0];
      res += vc[4 * vol + off] * xp[1];
      res += vc[5 * vol + off] * xp[g.px];
      if (DIM == 3) res += vc[6 * vol + off] * xp[pl];
      diag = (double)vc[3 * vol + off];
    } else {
      for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int e = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
            res += vc[e * vol + off] * xp[dz * pl + dy * g.px + dx];
          }
      diag = (double)vc[13 * vol + off];
    }
  } else {
#pragma unroll
    for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const double a = st.c[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
          if (a != 0.0) res += (T)a * xp[dz * pl + dy * g.px + dx];
        }
  }
  T c = f[chain * g.fcs + off];
  if (NOISE == NOISE_PHILOX) {
    float xi[4];
    const int ngx = g.px >> 3;
    const uint32_t c0 = (uint32_t)((((long long)k * g.py + j) * ngx + (i >> 3)) << 1) | (uint32_t)(i & 1);
    normals4(c0, (uint32_t)chain, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
    c += (T)(vc ? __builtin_sqrt(diag * (2.0 - st.omega) / st.omega) : st.noise_scale) * (T)xi[(i & 7) >> 1];
  }
  if (NOISE == NOISE_EXTERNAL)
    c += (T)(vc ? __builtin_sqrt(diag * (2.0 - st.omega) / st.omega) : st.noise_scale) * ext[chain * g.cs + off];
  *xp += (T)st.omega * (c - res) / (T)diag;
}

// ---------------------------------------------------------------------------
// Low-rank (measurement) terms, in two kernels.
//
// k_lr_dot: block (k, c) forms z[c][k] = sum_e B_ek x_c[idx_e] (B^T x, FP64
// accumulation); k_lr_solve: u_c = M (a sqrt(Sigma) xi'_c + zc z_c) (m x m,
// FP64, one block per chain); k_lr_update: x_c[row_r] += sum_e val_e u_c[col_e].
//
// Smoother:   M = S^{-1}, a = 0, zc = -1  ->  x -= W S^{-1} B^T x = B_bar B^T x
//             (src/smoother/sor_smoother.cc:46-51).
// Sampler:    M = S^{-1}, a = 1, zc = -1  ->  the SMW step with the sampler's
//             low-rank noise B Sigma^{-1/2} xi' folded in
//             (src/sampler/sor_sampler.cc:50-56): x' = y + W S^{-1} (Sigma^{1/2} xi' - B^T y).
// Residual:   M = Sigma^{-1}, a = 0, zc = -1 -> subtract B Sigma^{-1} B^T x.
// ---------------------------------------------------------------------------
constexpr int MMAX = 128;
// z[c][k] = (B^T x_c)_k for the LR_CPB (1, 2, 4) chains c0 .. c0 + LR_CPB - 1 of a
// block (grid: columns x chain groups): the column's indices and values are
// read once and the chains' x loads are independent.
template <typename T, int LR_CPB>
__device__ __forceinline__ void b_lr_dot(const Blk bk, const T* __restrict__ x, long long cs, const int32_t* __restrict__ bptr,
                                                const long long* __restrict__ bidx, const double* __restrict__ bval,
                                                int m, int K, double* __restrict__ z, int nsplit) {
  // block (kcol * nsplit + part, chain group): part `part` of nsplit equal
  // runs of the column's entries; the partial sums are added in order by
  // the solve
  const int kcol = bk.x / nsplit, part = bk.x % nsplit, c0 = bk.y * LR_CPB;
  const int nc = K - c0 < LR_CPB ? K - c0 : LR_CPB;
  double s[LR_CPB];
#pragma unroll
  for (int q = 0; q < LR_CPB; ++q) s[q] = 0.0;
  const int cb0 = bptr[kcol], clen = bptr[kcol + 1] - cb0;
  const int e1 = cb0 + (int)(((long long)clen * (part + 1)) / nsplit);
  const int stride = blockDim.x;
  int e = cb0 + (int)(((long long)clen * part) / nsplit) + threadIdx.x;
  // four entries per trip: their index and value loads, then their x loads,
  // are issued together (latency-bound gather)
  for (; e + 3 * stride < e1; e += 4 * stride) {
    long long ix[4];
    double bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ix[u] = bidx[e + u * stride];
      bv[u] = bval[e + u * stride];
    }
    T xv[4][LR_CPB];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < LR_CPB; ++q) xv[u][q] = q < nc ? x[(long long)(c0 + q) * cs + ix[u]] : (T)0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < LR_CPB; ++q) s[q] += bv[u] * (double)xv[u][q];
  }
  for (; e < e1; e += stride) {
    const double b = bval[e];
    const T* xp = x + (long long)c0 * cs + bidx[e];
#pragma unroll
    for (int q = 0; q < LR_CPB; ++q)
      if (q < nc) s[q] += b * (double)xp[q * cs];
  }
  __shared__ double red[LR_CPB][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < LR_CPB; ++q) {
    double v = s[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    const int q = threadIdx.x;
    z[((long long)(c0 + q) * m + kcol) * nsplit + part] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
  }
}
template <typename T, int LR_CPB>
__global__ void __launch_bounds__(256) k_lr_dot(const T* __restrict__ x, long long cs, const int32_t* __restrict__ bptr,
                                                const long long* __restrict__ bidx, const double* __restrict__ bval,
                                                int m, int K, double* __restrict__ z, int nsplit) {
  b_lr_dot<T, LR_CPB>(Blk::hw(), x, cs, bptr, bidx, bval, m, K, z, nsplit);
}
template <typename T, int LR_CPB>
constexpr Batchable<b_lr_dot<T, LR_CPB>, 256, false, k_lr_dot<T, LR_CPB>> kb_lr_dot{};

// B^T x for an x that is y + W u_p, the state right after an SMW step of
// the same level (y the sweep's output, z_p = B^T y the partial sums in z,
// u_p = M_p (sqrt(Sigma) xi'_p - z_p) the step's solve, W = T^{-1} B):
// B^T x = z_p + G u_p with G = B^T W (m x m), so the residual's low-rank term
// needs no second pass over x. G = nullptr: z holds B^T x itself.
struct LrCarry {
  const double* G;
  const double* Mp;    // the step's M
  int noise;           // its low-rank noise (NOISE_*), key and external values
  NoiseKey nk;
  const double* ext;
};

// sqrt(Sigma_i) xi'_i of chain c: the low-rank noise of an SMW sampler step
__device__ __forceinline__ double lr_noise(int noise, int i, int c, const NoiseKey& nk,
                                           const double* __restrict__ sqrt_sigma, const double* __restrict__ ext, int m) {
  if (noise == NOISE_PHILOX) {
    float xi[4];
    normals4(LOWRANK_SLOT + (uint32_t)(i >> 2), (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
    return sqrt_sigma[i] * (double)xi[i & 3];
  }
  if (noise == NOISE_EXTERNAL) return sqrt_sigma[i] * ext[(long long)c * m + i];
  return 0.0;
}

// one block (MMAX threads) per chain: u_c = M (a sqrt(Sigma) xi'_c + zc z_c),
// z_c from the partial sums, or with a carry z_c = z_p + G u_p (LrCarry)
template <int NOISE>
__device__ __forceinline__ void b_lr_solve(const Blk bk, const double* __restrict__ z, const double* __restrict__ M,
                                                   const double* __restrict__ sqrt_sigma, double zc,
                                                   const double* __restrict__ ext_lr, double* __restrict__ u, int m,
                                                   NoiseKey nk, int nsplit, LrCarry cy) {
  __shared__ double a[MMAX], t[MMAX];
  const int c = bk.x, tid = threadIdx.x;
  double zs = 0.0;
  if (tid < m)
    for (int p = 0; p < nsplit; ++p) zs += z[((long long)c * m + tid) * nsplit + p];
  if (cy.G) {
    // u_p as the step's own solve formed it, then z_p + G u_p
    if (tid < m) a[tid] = -zs + lr_noise(cy.noise, tid, c, cy.nk, sqrt_sigma, cy.ext, m);
    __syncthreads();
    if (tid < m) {
      double acc = 0.0;
#pragma unroll 8
      for (int l = 0; l < m; ++l) acc += cy.Mp[tid * m + l] * a[l];
      t[tid] = acc;
    }
    __syncthreads();
    if (tid < m) {
      double acc = zs;
#pragma unroll 8
      for (int l = 0; l < m; ++l) acc += cy.G[tid * m + l] * t[l];
      zs = acc;
    }
    __syncthreads();  // every a[l] is read before it is overwritten
  }
  if (tid < m) {
    double v = zc * zs;
    if (NOISE == NOISE_PHILOX) {
      float xi[4];
      normals4(LOWRANK_SLOT + (uint32_t)(tid >> 2), (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
      v += sqrt_sigma[tid] * (double)xi[tid & 3];
    }
    if (NOISE == NOISE_EXTERNAL) v += sqrt_sigma[tid] * ext_lr[(long long)c * m + tid];
    a[tid] = v;
  }
  __syncthreads();
  if (tid < m) {
    double acc = 0.0;
#pragma unroll 8
    for (int l = 0; l < m; ++l) acc += M[tid * m + l] * a[l];
    u[(long long)c * m + tid] = acc;
  }
}
template <int NOISE>
__global__ void __launch_bounds__(MMAX) k_lr_solve(const double* __restrict__ z, const double* __restrict__ M,
                                                   const double* __restrict__ sqrt_sigma, double zc,
                                                   const double* __restrict__ ext_lr, double* __restrict__ u, int m,
                                                   NoiseKey nk, int nsplit, LrCarry cy) {
  b_lr_solve<NOISE>(Blk::hw(), z, M, sqrt_sigma, zc, ext_lr, u, m, nk, nsplit, cy);
}
template <int NOISE>
constexpr Batchable<b_lr_solve<NOISE>, MMAX, false, k_lr_solve<NOISE>> kb_lr_solve{};

// x_c[rows[r]] += (R u_c)_r for the CPU chains of a block (grid: row
// blocks x chain groups); a row's entries are read once for all its chains.
constexpr int LR_CPU = 8;
template <typename T, int CPU>
__device__ __forceinline__ void b_lr_update(const Blk bk, T* __restrict__ x, long long cs, const long long* __restrict__ rows,
                                                   const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                   const double* __restrict__ val, int nrows,
                                                   const double* __restrict__ u, int m, int K) {
  __shared__ double us[CPU][MMAX];
  const int c0 = bk.y * CPU;
  const int nc = K - c0 < CPU ? K - c0 : CPU;
  const int tid = threadIdx.x;
  const int r = bk.x * blockDim.x + tid;
  const bool ok = r < nrows;
  // the row's x values and W entries are requested before the block waits
  // for u in LDS (latency-bound scatter)
  long long rr = 0;
  int e0 = 0, e1 = 0;
  if (ok) {
    rr = rows[r];
    e0 = ptr[r];
    e1 = ptr[r + 1];
  }
  T xv[CPU];
#pragma unroll
  for (int q = 0; q < CPU; ++q) xv[q] = (ok && q < nc) ? x[(long long)(c0 + q) * cs + rr] : (T)0;
  for (int e = tid; e < nc * m; e += blockDim.x) us[e / m][e % m] = u[(long long)c0 * m + e];
  __syncthreads();
  if (!ok) return;
  double s[CPU];
#pragma unroll
  for (int q = 0; q < CPU; ++q) s[q] = 0.0;
  for (int e = e0; e < e1; ++e) {
    const double v = val[e];
    const int cl = col[e];
#pragma unroll
    for (int q = 0; q < CPU; ++q) s[q] += v * us[q][cl];
  }
  T* p = x + (long long)c0 * cs + rr;
#pragma unroll
  for (int q = 0; q < CPU; ++q)
    if (q < nc) p[q * cs] = (T)((double)xv[q] + s[q]);
}
template <typename T, int CPU>
__global__ void __launch_bounds__(256) k_lr_update(T* __restrict__ x, long long cs, const long long* __restrict__ rows,
                                                   const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                   const double* __restrict__ val, int nrows,
                                                   const double* __restrict__ u, int m, int K) {
  b_lr_update<T, CPU>(Blk::hw(), x, cs, rows, ptr, col, val, nrows, u, m, K);
}
template <typename T, int CPU>
constexpr Batchable<b_lr_update<T, CPU>, 256, false, k_lr_update<T, CPU>> kb_lr_update{};

// k_lr_solve and k_lr_update in one launch: every block first forms u for
// its LR_CPU chains (u_c = M (zc z_c + sqrt(Sigma) xi'_c), the same sums in
// the same order as k_lr_solve, recomputed per row block: m^2 FMAs per chain),
// then applies x_c[rows[r]] += (R u_c)_r.
template <typename T, int NOISE>
__device__ __forceinline__ void b_lr_solve_update(const Blk bk, T* __restrict__ x, long long cs,
                                                         const long long* __restrict__ rows,
                                                         const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ col,
                                                         const double* __restrict__ val, int nrows,
                                                         const double* __restrict__ z, const double* __restrict__ M,
                                                         const double* __restrict__ sqrt_sigma, double zc,
                                                         const double* __restrict__ ext_lr, int m, int K, NoiseKey nk,
                                                         int nsplit, LrCarry cy) {
  __shared__ double a[LR_CPU][MMAX];
  __shared__ double us[LR_CPU][MMAX];
  const int c0 = bk.y * LR_CPU;
  const int nc = K - c0 < LR_CPU ? K - c0 : LR_CPU;
  const int tid = threadIdx.x;
  const int r = bk.x * blockDim.x + tid;
  const bool ok = r < nrows;
  long long rr = 0;
  int e0 = 0, e1 = 0;
  if (ok) {
    rr = rows[r];
    e0 = ptr[r];
    e1 = ptr[r + 1];
  }
  T xv[LR_CPU];
#pragma unroll
  for (int q = 0; q < LR_CPU; ++q) xv[q] = (ok && q < nc) ? x[(long long)(c0 + q) * cs + rr] : (T)0;
  if (cy.G) {
    // with a carry (LrCarry): u_p of the SMW step before, in us
    for (int e = tid; e < nc * m; e += blockDim.x) {
      const int q = e / m, i = e % m, c = c0 + q;
      double zs = 0.0;
      for (int p = 0; p < nsplit; ++p) zs += z[((long long)c * m + i) * nsplit + p];
      a[q][i] = -zs + lr_noise(cy.noise, i, c, cy.nk, sqrt_sigma, cy.ext, m);
    }
    __syncthreads();
    for (int e = tid; e < nc * m; e += blockDim.x) {
      const int q = e / m, i = e % m;
      double acc = 0.0;
#pragma unroll 8
      for (int l = 0; l < m; ++l) acc += cy.Mp[i * m + l] * a[q][l];
      us[q][i] = acc;
    }
    __syncthreads();
  }
  for (int e = tid; e < nc * m; e += blockDim.x) {
    const int q = e / m, i = e % m, c = c0 + q;
    double zs = 0.0;
    for (int p = 0; p < nsplit; ++p) zs += z[((long long)c * m + i) * nsplit + p];
    if (cy.G) {
      double acc = zs;
#pragma unroll 8
      for (int l = 0; l < m; ++l) acc += cy.G[i * m + l] * us[q][l];
      zs = acc;
    }
    double v = zc * zs;
    if (NOISE == NOISE_PHILOX) {
      float xi[4];
      normals4(LOWRANK_SLOT + (uint32_t)(i >> 2), (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
      v += sqrt_sigma[i] * (double)xi[i & 3];
    }
    if (NOISE == NOISE_EXTERNAL) v += sqrt_sigma[i] * ext_lr[(long long)c * m + i];
    a[q][i] = v;
  }
  __syncthreads();
  for (int e = tid; e < nc * m; e += blockDim.x) {
    const int q = e / m, i = e % m;
    const double* Mi = M + i * m;
    double acc = 0.0;
#pragma unroll 8
    for (int l = 0; l < m; ++l) acc += Mi[l] * a[q][l];
    us[q][i] = acc;
  }
  __syncthreads();
  if (!ok) return;
  double s[LR_CPU];
#pragma unroll
  for (int q = 0; q < LR_CPU; ++q) s[q] = 0.0;
  for (int e = e0; e < e1; ++e) {
    const double v = val[e];
    const int cl = col[e];
#pragma unroll
    for (int q = 0; q < LR_CPU; ++q) s[q] += v * us[q][cl];
  }
  T* p = x + (long long)c0 * cs + rr;
#pragma unroll
  for (int q = 0; q < LR_CPU; ++q)
    if (q < nc) p[q * cs] = (T)((double)xv[q] + s[q]);
}
template <typename T, int NOISE>
__global__ void __launch_bounds__(256) k_lr_solve_update(T* __restrict__ x, long long cs,
                                                         const long long* __restrict__ rows,
                                                         const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ col,
                                                         const double* __restrict__ val, int nrows,
                                                         const double* __restrict__ z, const double* __restrict__ M,
                                                         const double* __restrict__ sqrt_sigma, double zc,
                                                         const double* __restrict__ ext_lr, int m, int K, NoiseKey nk,
                                                         int nsplit, LrCarry cy) {
  b_lr_solve_update<T, NOISE>(Blk::hw(), x, cs, rows, ptr, col, val, nrows, z, M, sqrt_sigma, zc, ext_lr, m, K, nk, nsplit,
                              cy);
}
template <typename T, int NOISE>
constexpr Batchable<b_lr_solve_update<T, NOISE>, 256, false, k_lr_solve_update<T, NOISE>> kb_lr_solve_update{};

// ---------------------------------------------------------------------------
// Transfers. Grids are (groups x rows, chains): thread t of chain
// blockIdx.y owns group t % ng of interior row t / ng, row id
// (j - 1) + (n1 - 1)(k - 1); an aligned 8-vertex group (vector loads and
// stores, 32-bit index arithmetic only).
// ---------------------------------------------------------------------------

// Row-block grids (x: blocks of rows, y: chains). Workgroups are dispatched
// to the eight XCDs round-robin in linear order (x fastest), so when the grid
// divides by 8 the (block, chain) items are dealt out such that an XCD runs
// consecutive row blocks of one chain, whose shared neighbour rows and
// planes then meet in its L2.
__device__ __forceinline__ void xcd_rowblock(const Blk& bk, int& bx, int& c) {
  const int gxn = (int)bk.gx, nb = (int)(bk.gx * bk.gy);
  int lin = (int)bk.x + (int)bk.y * gxn;
  if ((nb & 7) == 0) lin = (lin & 7) * (nb >> 3) + (lin >> 3);
  bx = lin % gxn;
  c = lin / gxn;
}

// r = f - A x on interior vertices (the sparse part of
// src/linear_operator/linear_operator.hh:78-88, multigridmc_sampler.cc:125-126)
template <typename T, int DIM, bool STAR>
__device__ __forceinline__ void b_residual(const Blk bk, const T* __restrict__ x, const T* __restrict__ f, T* __restrict__ r,
                                                  DevGeom g, Stencil st) {
  const int ngx = g.px >> 3;
  const int nrows = (g.n1 - 1) * (DIM == 3 ? g.n2 - 1 : 1);
  int bx, c;
  xcd_rowblock(bk, bx, c);
  const int t = bx * blockDim.x + threadIdx.x;
  if (t >= ngx * nrows) return;
  const int gx = t % ngx, row = t / ngx;
  const int j = row % (g.n1 - 1) + 1;
  const int k = DIM == 3 ? row / (g.n1 - 1) + 1 : 0;
  const long long pl = (long long)g.px * g.py;
  const long long base = (long long)c * g.cs + ((long long)k * g.py + j) * g.px + 8 * gx;
  const T* xr = x + base;
  T xc[10];
  load8<T>(xr, xc + 1);
  xc[0] = xr[-1];
  xc[9] = xr[8];
  T acc[8];
  if (STAR) {
    T ym[8], yp[8];
    load8<T>(xr - g.px, ym);
    load8<T>(xr + g.px, yp);
    const T cx = (T)st.c[12], cy = (T)st.c[10], D = (T)st.diag;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = cy * ym[q] + cx * xc[q] + D * xc[q + 1] + cx * xc[q + 2] + cy * yp[q];
    if (DIM == 3) {
      T zm[8], zp[8];
      load8<T>(xr - pl, zm);
      load8<T>(xr + pl, zp);
      const T cz = (T)st.c[4];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = cz * zm[q] + acc[q] + cz * zp[q];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = (T)0;
#pragma unroll
    for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        T w[10];
        const T* p = xr + dz * pl + dy * g.px;
        load8<T>(p, w + 1);
        w[0] = p[-1];
        w[9] = p[8];
        const T a0 = (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + 0];
        const T a1 = (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + 1];
        const T a2 = (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + 2];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] += a0 * w[q] + a1 * w[q + 1] + a2 * w[q + 2];
      }
  }
  T fb[8], out[8];
  if (f) load8<T>(f + (long long)c * g.fcs + ((long long)k * g.py + j) * g.px + 8 * gx, fb);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 8 * gx + q;
    out[q] = (i >= 1 && i <= g.n0 - 1) ? (f ? fb[q] : (T)0) - acc[q] : (T)0;
  }
  store8<T>(r + base, out);
}
template <typename T, int DIM, bool STAR>
__global__ void __launch_bounds__(256) k_residual(const T* __restrict__ x, const T* __restrict__ f, T* __restrict__ r,
                                                  DevGeom g, Stencil st) {
  b_residual<T, DIM, STAR>(Blk::hw(), x, f, r, g, st);
}
template <typename T, int DIM, bool STAR>
constexpr Batchable<b_residual<T, DIM, STAR>, 256, false, k_residual<T, DIM, STAR>> kb_residual{};

// fc[I] = sum_sigma p(sigma) r[2I + sigma], p = prod {1/2, 1, 1/2}
// (IntergridOperator::restrict, src/intergrid/intergrid_operator.hh:74-88).
// A thread owns four coarse vertices I = 4 gq .. 4 gq + 3 of a coarse row,
// i.e. fine x positions 8 gq - 1 .. 8 gq + 7.
template <typename T, int DIM>
__device__ __forceinline__ void b_restrict(const Blk bk, const T* __restrict__ r, T* __restrict__ fc, DevGeom gf, DevGeom gc) {
  const int nq = (gc.n0 - 1 + 4) / 4;  // groups of four coarse vertices covering I = 0 .. n0c - 1
  const int nrows = (gc.n1 - 1) * (DIM == 3 ? gc.n2 - 1 : 1);
  int bx, c;
  xcd_rowblock(bk, bx, c);
  const int t = bx * blockDim.x + threadIdx.x;
  if (t >= nq * nrows) return;
  const int gq = t % nq, row = t / nq;
  const int J = row % (gc.n1 - 1) + 1;
  const int Kz = DIM == 3 ? row / (gc.n1 - 1) + 1 : 0;
  const long long pl = (long long)gf.px * gf.py;
  const long long fb0 = (long long)c * gf.cs + ((long long)(2 * Kz) * gf.py + 2 * J) * gf.px + 8 * gq;
  const T w[3] = {(T)0.5, (T)1.0, (T)0.5};
  T acc[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
  for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const T* q = r + fb0 + dz * pl + dy * gf.px;
      T v[10];
      load8<T>(q, v + 1);
      v[0] = q[-1];
      v[9] = q[8];
      const T wzy = (DIM == 3 ? w[dz + 1] : (T)1) * w[dy + 1];
#pragma unroll
      for (int a = 0; a < 4; ++a) acc[a] += wzy * ((T)0.5 * v[2 * a] + v[2 * a + 1] + (T)0.5 * v[2 * a + 2]);
    }
  T* out = fc + (long long)c * gc.cs + ((long long)Kz * gc.py + J) * gc.px + 4 * gq;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int I = 4 * gq + a;
    if (I >= 1 && I <= gc.n0 - 1) out[a] = acc[a];
  }
}
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_restrict(const T* __restrict__ r, T* __restrict__ fc, DevGeom gf, DevGeom gc) {
  b_restrict<T, DIM>(Blk::hw(), r, fc, gf, gc);
}
template <typename T, int DIM>
constexpr Batchable<b_restrict<T, DIM>, 256, false, k_restrict<T, DIM>> kb_restrict{};

// x += alpha P xc (IntergridOperator::prolongate_add in gather form,
// src/intergrid/intergrid_operator.hh:106-120). A thread owns an aligned
// 8-vertex fine group, whose coarse neighbours are I = 4 gx .. 4 gx + 4.
template <typename T, int DIM>
__device__ __forceinline__ void b_prolongate(const Blk bk, T* __restrict__ x, const T* __restrict__ xc, DevGeom gf, DevGeom gc,
                                                    T alpha) {
  const int ngx = gf.px >> 3;
  const int nrows = (gf.n1 - 1) * (DIM == 3 ? gf.n2 - 1 : 1);
  int bx, c;
  xcd_rowblock(bk, bx, c);
  const int t = bx * blockDim.x + threadIdx.x;
  if (t >= ngx * nrows) return;
  const int gx = t % ngx, row = t / ngx;
  const int j = row % (gf.n1 - 1) + 1;
  const int k = DIM == 3 ? row / (gf.n1 - 1) + 1 : 0;
  // coarse rows/planes touching fine row j / plane k, with weights
  const int j0 = j >> 1, j1 = (j + 1) >> 1;
  const int k0 = k >> 1, k1 = (k + 1) >> 1;
  const T half = (T)0.5;
  const T* cb = xc + (long long)c * gc.cs;
  T s[5] = {(T)0, (T)0, (T)0, (T)0, (T)0};  // y/z-interpolated coarse values at I = 4 gx .. 4 gx + 4
  const int nz = (DIM == 3 && (k & 1)) ? 2 : 1, ny = (j & 1) ? 2 : 1;
  const T wz = nz == 2 ? half : (T)1, wy = ny == 2 ? half : (T)1;
  for (int a = 0; a < nz; ++a)
    for (int b2 = 0; b2 < ny; ++b2) {
      const int K = a ? k1 : k0, J = b2 ? j1 : j0;
      const T* p = cb + ((long long)K * gc.py + J) * gc.px + 4 * gx;
#pragma unroll
      for (int e = 0; e < 5; ++e) s[e] += (4 * gx + e < gc.px) ? p[e] : (T)0;
    }
  const T w = wz * wy * alpha;
  T* xr = x + (long long)c * gf.cs + ((long long)k * gf.py + j) * gf.px + 8 * gx;
  T v[8];
  load8<T>(xr, v);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 8 * gx + q;
    const T p = (q & 1) ? half * (s[q >> 1] + s[(q >> 1) + 1]) : s[q >> 1];
    if (i >= 1 && i <= gf.n0 - 1) v[q] += w * p;
  }
  store8<T>(xr, v);
}
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_prolongate(T* __restrict__ x, const T* __restrict__ xc, DevGeom gf, DevGeom gc,
                                                    T alpha) {
  b_prolongate<T, DIM>(Blk::hw(), x, xc, gf, gc, alpha);
}
template <typename T, int DIM>
constexpr Batchable<b_prolongate<T, DIM>, 256, false, k_prolongate<T, DIM>> kb_prolongate{};

template <typename T>
__device__ __forceinline__ void b_copy(const Blk bk, T* __restrict__ dst, const T* __restrict__ src, long long n) {
  long long t = (long long)bk.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}
template <typename T>
__global__ void __launch_bounds__(256) k_copy(T* __restrict__ dst, const T* __restrict__ src, long long n) {
  b_copy<T>(Blk::hw(), dst, src, n);
}
template <typename T>
constexpr Batchable<b_copy<T>, 256, false, k_copy<T>> kb_copy{};

template <typename T>
__global__ void __launch_bounds__(256) k_add(T* __restrict__ dst, const T* __restrict__ src, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] += src[t];
}

template <typename T>
__device__ __forceinline__ void b_fill(const Blk bk, T* __restrict__ p, long long n, T v) {
  long long t = (long long)bk.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ p, long long n, T v) {
  b_fill<T>(Blk::hw(), p, n, v);
}
template <typename T>
constexpr Batchable<b_fill<T>, 256, false, k_fill<T>> kb_fill{};

// reference order (interior, x fastest, FP64) -> padded level layout
template <typename T>
__global__ void __launch_bounds__(256) k_pack(const double* __restrict__ ref, T* __restrict__ pad, DevGeom g, int K,
                                              int per_chain) {
  long long per = (long long)(g.n0 - 1) * (g.n1 - 1) * (g.dim == 3 ? g.n2 - 1 : 1);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * K) return;
  int c = (int)(t / per);
  long long e = t % per;
  long long src = per_chain ? t : e;
  int i = (int)(e % (g.n0 - 1)) + 1;
  e /= (g.n0 - 1);
  int j = (int)(e % (g.n1 - 1)) + 1;
  int k = g.dim == 3 ? (int)(e / (g.n1 - 1)) + 1 : 0;
  pad[c * g.cs + ((long long)k * g.py + j) * g.px + i] = (T)ref[src];
}

template <typename T>
__global__ void __launch_bounds__(256) k_unpack(const T* __restrict__ pad, double* __restrict__ ref, DevGeom g, int K,
                                                double scale) {
  long long per = (long long)(g.n0 - 1) * (g.n1 - 1) * (g.dim == 3 ? g.n2 - 1 : 1);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * K) return;
  int c = (int)(t / per);
  long long e = t % per;
  int i = (int)(e % (g.n0 - 1)) + 1;
  e /= (g.n0 - 1);
  int j = (int)(e % (g.n1 - 1)) + 1;
  int k = g.dim == 3 ? (int)(e / (g.n1 - 1)) + 1 : 0;
  ref[t] = scale * (double)pad[c * g.cs + ((long long)k * g.py + j) * g.px + i];
}

// Coarsest level: x = L^{-T} xi + Q^{-1} f (= L^{-T}(xi + L^{-1} f),
// src/sampler/cholesky_sampler.hh:52-66) or x = Q^{-1} f
// (src/solver/cholesky_solver.cc:30-40); FP64 inside. Grid (row blocks of 64,
// groups of CPT chains); the P = blockDim / 64 waves of a block split the
// inner dimension and reduce through LDS in a fixed order, and every matrix
// element a thread loads serves the CPT chains of the block (right-hand
// sides and normals staged in LDS). The matrices are stored column-major
// (element (r, j) at j * Nc + r) so that a wave reads 64 consecutive rows of
// one column. LDS: max(2 CPT Nc, P CPT 64) doubles.
template <typename T, int NOISE, int CPT>
__device__ __forceinline__ void b_coarse(const Blk bk, T* __restrict__ x, const T* __restrict__ f, long long cs, long long fcs,
                                                 const long long* __restrict__ pidx, const double* __restrict__ LT,
                                                 const double* __restrict__ Q, int Nc, int K,
                                                 const double* __restrict__ ext, NoiseKey nk) {
  MGMC_DYNAMIC_SHARED(double, sh);
  double* fs = sh;             // [CPT][Nc]
  double* xs = sh + CPT * Nc;  // [CPT][Nc]
  const int c0 = bk.y * CPT;
  const int tid = threadIdx.x;
  for (int e = tid; e < CPT * Nc; e += blockDim.x) {
    const int q = e / Nc, j = e % Nc, c = c0 + q;
    double fv = 0.0, xv = 0.0;
    if (c < K) {
      fv = (double)f[c * fcs + pidx[j]];
      if (NOISE == NOISE_PHILOX) {
        float xi[4];
        normals4(COARSE_SLOT + (uint32_t)(j >> 2), (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
        xv = (double)xi[j & 3];
      } else if (NOISE == NOISE_EXTERNAL) {
        xv = ext[(long long)c * Nc + j];
      }
    }
    fs[e] = fv;
    xs[e] = xv;
  }
  __syncthreads();
  const int P = (int)blockDim.x >> 6;
  const int lane = tid & 63, part = tid >> 6;
  const int r = bk.x * 64 + lane;
  const int j0 = (Nc * part) / P, j1 = (Nc * (part + 1)) / P;
  double s[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) s[c] = 0.0;
  if (r < Nc) {
    for (int j = j0; j < j1; ++j) {
      const double q = Q[(long long)j * Nc + r];
#pragma unroll
      for (int c = 0; c < CPT; ++c) s[c] += q * fs[c * Nc + j];
    }
    if (NOISE != NOISE_NONE)
      for (int j = j0 > r ? j0 : r; j < j1; ++j) {
        const double l = LT[(long long)j * Nc + r];
#pragma unroll
        for (int c = 0; c < CPT; ++c) s[c] += l * xs[c * Nc + j];
      }
  }
  __syncthreads();  // fs and xs are consumed: the same LDS holds the partial sums
  double* red = sh;  // [P][CPT][64]
#pragma unroll
  for (int c = 0; c < CPT; ++c) red[(part * CPT + c) * 64 + lane] = s[c];
  __syncthreads();
  if (part == 0 && r < Nc) {
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      if (c0 + c >= K) break;
      double v = 0.0;
      for (int p = 0; p < P; ++p) v += red[(p * CPT + c) * 64 + lane];
      x[(c0 + c) * cs + pidx[r]] = (T)v;
    }
  }
}
template <typename T, int NOISE, int CPT>
__global__ void __launch_bounds__(1024) k_coarse(T* __restrict__ x, const T* __restrict__ f, long long cs, long long fcs,
                                                 const long long* __restrict__ pidx, const double* __restrict__ LT,
                                                 const double* __restrict__ Q, int Nc, int K,
                                                 const double* __restrict__ ext, NoiseKey nk) {
  b_coarse<T, NOISE, CPT>(Blk::hw(), x, f, cs, fcs, pidx, LT, Q, Nc, K, ext, nk);
}
template <typename T, int NOISE, int CPT>
constexpr Batchable<b_coarse<T, NOISE, CPT>, 1024, false, k_coarse<T, NOISE, CPT>> kb_coarse{};

// The same coarsest-level products on the FP64 matrix cores. A workgroup
// owns one 16 x 16 output tile (16 rows r0 .. r0 + 15 x 16 chains c0 ..),
// its waves split the inner dimension into equal runs of k-steps of four
// and reduce their partial tiles through LDS in a fixed order; each k-step
// is one v_mfma_f64_16x16x4f64 (lane l holds A[l % 16][l / 16] of the
// 16 x 4 operand and B[l / 16][l % 16] of the 4 x 16 one; item v of the
// accumulator is D[l / 16 + 4 v][l % 16]). The workgroup stages the
// right-hand sides and normals of its 16 chains in LDS (zero beyond K and
// beyond Nc); L^{-T} is upper triangular, so its steps below row r0 are
// skipped. LDS: 2 * 16 * round_up(Nc, 4) + 256 * waves doubles.
#ifndef MGMC_EMULATION
typedef double mgmc_d4 __attribute__((ext_vector_type(4)));
template <typename T, int NOISE>
__device__ __forceinline__ void b_coarse_mfma(const Blk bk, T* __restrict__ x, const T* __restrict__ f, long long cs,
                                                     long long fcs, const long long* __restrict__ pidx,
                                                     const double* __restrict__ LT, const double* __restrict__ Q,
                                                     int Nc, int K, const double* __restrict__ ext, NoiseKey nk) {
  MGMC_DYNAMIC_SHARED(double, sh);
  const int Nc4 = (Nc + 3) & ~3;
  double* fs = sh;                  // [16][Nc4]
  double* xs = sh + 16 * Nc4;       // [16][Nc4]
  double* red = sh + 32 * Nc4;      // [waves][256]
  const int r0 = bk.x * 16, c0 = bk.y * 16;
  const int tid = threadIdx.x, nw = (int)blockDim.x >> 6, w = tid >> 6, lane = tid & 63;
  for (int e = tid; e < 16 * Nc4; e += blockDim.x) {
    const int q = e / Nc4, j = e % Nc4, c = c0 + q;
    double fv = 0.0, xv = 0.0;
    if (c < K && j < Nc) {
      fv = (double)f[c * fcs + pidx[j]];
      if (NOISE == NOISE_PHILOX) {
        float xi[4];
        normals4(COARSE_SLOT + (uint32_t)(j >> 2), (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
        xv = (double)xi[j & 3];
      } else if (NOISE == NOISE_EXTERNAL) {
        xv = ext[(long long)c * Nc + j];
      }
    }
    fs[e] = fv;
    xs[e] = xv;
  }
  __syncthreads();
  const int ra = r0 + (lane & 15), ka = lane >> 4, cb = lane & 15;
  // this wave's k-steps: [s0, s1) of the Nc4 / 4 steps; four accumulators
  // (four matrix-core chains in flight), operand loads of a trip issued
  // together
  const int nstep = Nc4 >> 2;
  const int per = (nstep + nw - 1) / nw;
  const int s0 = w * per, s1 = s0 + per < nstep ? s0 + per : nstep;
  mgmc_d4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = mgmc_d4{0.0, 0.0, 0.0, 0.0};
  auto steps = [&](const double* __restrict__ M, const double* B, int sbeg, bool upper) {
    for (int st = sbeg; st < s1; st += 4) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int jj = 4 * (st + u) + ka;
        const bool in = st + u < s1;
        const bool ok = in && ra < Nc && jj < Nc && (!upper || jj >= ra);
        a[u] = ok ? M[(long long)jj * Nc + ra] : 0.0;
        b[u] = in ? B[cb * Nc4 + jj] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc[u], 0, 0, 0);
    }
  };
  steps(Q, fs, s0, false);
  if (NOISE != NOISE_NONE) steps(LT, xs, s0 > (r0 >> 2) ? s0 : (r0 >> 2), true);
#pragma unroll
  for (int v = 0; v < 4; ++v) red[w * 256 + v * 64 + lane] = ((acc[0][v] + acc[1][v]) + acc[2][v]) + acc[3][v];
  __syncthreads();
  if (w == 0) {
    const int c = c0 + cb;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double t = 0.0;
      for (int q = 0; q < nw; ++q) t += red[q * 256 + v * 64 + lane];
      const int r = r0 + ka + 4 * v;
      if (r < Nc && c < K) x[c * cs + pidx[r]] = (T)t;
    }
  }
}
template <typename T, int NOISE>
__global__ void __launch_bounds__(512) k_coarse_mfma(T* __restrict__ x, const T* __restrict__ f, long long cs,
                                                     long long fcs, const long long* __restrict__ pidx,
                                                     const double* __restrict__ LT, const double* __restrict__ Q,
                                                     int Nc, int K, const double* __restrict__ ext, NoiseKey nk) {
  b_coarse_mfma<T, NOISE>(Blk::hw(), x, f, cs, fcs, pidx, LT, Q, Nc, K, ext, nk);
}
template <typename T, int NOISE>
constexpr Batchable<b_coarse_mfma<T, NOISE>, 512, false, k_coarse_mfma<T, NOISE>> kb_coarse_mfma{};
#endif

// ---------------------------------------------------------------------------
// One whole red-black SOR/Gibbs sweep of a 3D star-stencil level in a single
// pass over HBM. A workgroup owns full x-rows (one thread per aligned group
// of eight vertices) for TY owned y-rows plus two halo rows on each side, and
// marches through its z-segment [ks, ke): at step k it applies the first
// colour to plane k and the second colour to plane k-1, then stores plane
// k-1. The first colour is also applied on the halo ring (rows j0-1, j0+TY,
// planes ks-1, ke) and discarded: those values depend only on old values and
// on counter-based noise keyed by the vertex, so they equal the values the
// owning workgroup computes. The sweep is out of place (x -> xout) so that
// every workgroup reads old values in its halo; every interior vertex of xout
// is written exactly once. Result = the two colour-phase launches of k_sweep
// (same Philox slots, same update formula).
// Per step: the loads of plane k+1 and f(k) are issued first and the Philox
// normals of both colours are computed while they are in flight; LDS holds
// plane k before the first colour (TY+4 rows) and plane k-1 after it (TY+2
// rows), each double-buffered so one barrier per step suffices.
// ---------------------------------------------------------------------------
// FP64 workgroups are capped at 512 threads, which gives the compiler 256
// VGPRs for the plane window instead of spilling it.
// Chain-blocked FP32 workgroups (CB = 2) are capped at 512 threads for the
// same reason: the plane windows of two chains need up to 256 VGPRs.
template <typename T, int CB = 1>
constexpr int rb3_max_threads() {
  return sizeof(T) == 8 || CB > 1 ? 512 : 1024;
}

// Chain blocking (CB = 2, f shared by the chains: fcs = 0): a workgroup
// sweeps two consecutive chains of its tile and segment. Each thread keeps
// the plane window of both chains, f(k) is loaded once for both and one
// barrier per step serves both updates; each chain has its own LDS planes.
// Every chain gets the noise, arithmetic and update order of CB = 1. The
// per-chain code is a macro over separately named arrays, which the compiler
// keeps in VGPRs (arrays indexed by the chain went to scratch memory).
template <typename T, int NOISE, int CB>
__device__ __forceinline__ void b_sweep_rb3(const Blk bk, const T* __restrict__ x, T* __restrict__ xout,
                                                    const T* __restrict__ f, const T* __restrict__ ext, DevGeom g,
                                                    Stencil st, int first, int TY, int SZ, int ntile, int nseg,
                                                    int ncol, NoiseKey nk, const T* __restrict__ xcoarse, DevGeom gc,
                                                    T alpha) {
  static_assert(CB == 1 || CB == 2, "one or two chains per workgroup");
  MGMC_DYNAMIC_SHARED(char, lds_raw);
  // LDS per chain: [guard][A (k even)][guard][A (k odd)][guard][B (k even)][guard][B (k odd)][guard];
  // the zero guards stand for the x neighbours beyond the first and last row
  constexpr int LG = 8;
  const size_t asz = (size_t)(TY + 4) * g.px + LG, bsz = (size_t)(TY + 2) * g.px + LG;
  const size_t csz = 2 * asz + 2 * bsz + LG;  // one chain's LDS
  for (int i = threadIdx.x; i < 5 * LG * CB; i += blockDim.x) {
    const int cb = i / (5 * LG), t = i % (5 * LG);
    const int e = t / LG, o = t % LG;
    T* A0c = reinterpret_cast<T*>(lds_raw) + cb * csz + LG;
    T* gptr = e < 3 ? A0c - LG + e * asz : A0c + 2 * asz - LG + (e - 2) * bsz;
    gptr[o] = (T)0;
  }
  T* const A0 = reinterpret_cast<T*>(lds_raw) + LG;  // chain 0; chain 1 at A0 + csz
  const int ngx = g.px >> 3;
  const int tid = threadIdx.x;
  const int r = tid / ngx, gx = tid % ngx;
  const bool active = r < TY + 4;
  const long long pl = (long long)g.px * g.py;
  const uint32_t sample = *nk.sample;
  const int second = 1 - first;
  // groups holding no interior vertex (the x padding) draw no noise and update nothing
  const bool inner = 8 * gx <= g.n0 - 1;
  const T cx = (T)st.c[12], cy = (T)st.c[10], cz = (T)st.c[4], D = (T)st.diag;
  const T om_d = (T)st.inv_diag_omega, ns = (T)st.noise_scale;
  const size_t aoff = (size_t)r * g.px;        // A rows j0-2 .. j0+TY+1 (chunked, LdsRow)
  const size_t boff = (size_t)(r - 1) * g.px;  // B rows j0-1 .. j0+TY (r = 1 .. TY+2)

  // the march of chain(s) `chain` over tile `tile`, owned interior planes [ks, ke)
  auto march = [&](const int tile, const int chain, const int ks, const int ke) {
    const int j0 = 1 + tile * TY;
    const int j = j0 - 2 + r;
    const bool has_mem = active && j >= 0 && j <= g.n1;
    const bool ring = active && j >= j0 - 1 && j <= j0 + TY && j >= 1 && j <= g.n1 - 1;
    const bool owned = active && j >= j0 && j <= j0 + TY - 1 && j <= g.n1 - 1;
    const long long cbase = (long long)chain * g.cs + (long long)j * g.px + 8 * gx;
    const long long cbase1 = cbase + g.cs;  // second chain (CB = 2)
    const long long fbase = (long long)chain * g.fcs + (long long)j * g.px + 8 * gx;

    T xmm[8], xm[8], xc[8], xp[8], fm[8], fc[8];
    T xmm1[8], xm1[8], xc1[8], xp1[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) xmm[q] = xm[q] = xc[q] = xp[q] = fm[q] = fc[q] = xmm1[q] = xm1[q] = xc1[q] = xp1[q] = (T)0;
    // optional prolongate-add of the coarse correction while the planes are read
    ProlongStream<T> pro, pro1;
#define MGMC_RB3_INIT(XM, XC, PRO, CH, CBASE)                                                                    \
    PRO.init(xcoarse && has_mem ? xcoarse + (long long)(CH) * gc.cs : nullptr, gc, gx, j, ks - 2 >= 0 ? ks - 2 : ks - 1, \
             alpha);                                                                                                 \
    if (has_mem && ks - 2 >= 0) {                                                                                    \
      load8<T>(x + CBASE + (ks - 2) * pl, XM);                                                                       \
      PRO.add(XM, ks - 2);                                                                                           \
    }                                                                                                                \
    if (has_mem) {                                                                                                   \
      load8<T>(x + CBASE + (ks - 1) * pl, XC);                                                                       \
      PRO.add(XC, ks - 1);                                                                                           \
    }
    MGMC_RB3_INIT(xm, xc, pro, chain, cbase)
    if constexpr (CB > 1) {
      MGMC_RB3_INIT(xm1, xc1, pro1, chain + 1, cbase1)
    }
#undef MGMC_RB3_INIT

    // step k: xmm, xm, xc = planes k-2, k-1, k (xm after the first colour), xp
    // receives plane k+1; fm, fc = f on planes k-1, k
    for (int k = ks - 1; k <= ke; ++k) {
      const int kb = k - 1;
      const bool doA = ring && inner && k >= 1 && k <= g.n2 - 1;
      const bool storeB = owned && kb >= ks && kb <= ke - 1;
      const bool doB = storeB && inner;
      // loads first: next plane of x, f on plane k
      // (the coarse rows of plane k+1's correction are requested with it and
      // added after the noise, before the first colour reads plane k+1)
      const bool next = has_mem && k + 1 <= g.n2;
#define MGMC_RB3_LOAD(XP, PRO, CBASE)                     \
    if (next) {                                             \
      load8<T>(x + CBASE + (long long)(k + 1) * pl, XP);    \
      PRO.advance(k + 1);                                   \
    } else {                                                \
      _Pragma("unroll") for (int q = 0; q < 8; ++q) XP[q] = (T)0; \
    }
      MGMC_RB3_LOAD(xp, pro, cbase)
      if constexpr (CB > 1) {
        MGMC_RB3_LOAD(xp1, pro1, cbase1)
      }
#undef MGMC_RB3_LOAD
      if (doA) load8<T>(f + fbase + (long long)k * pl, fc);
      // noise of both colours while the loads are in flight
      const int iparA = (first + j + k) & 1, iparB = (second + j + kb) & 1;
      float xiA[4] = {0.f, 0.f, 0.f, 0.f}, xiB[4] = {0.f, 0.f, 0.f, 0.f};
      float xiA1[4] = {0.f, 0.f, 0.f, 0.f}, xiB1[4] = {0.f, 0.f, 0.f, 0.f};
      if (NOISE == NOISE_PHILOX) {
        const uint32_t cj = (uint32_t)(j * ngx + gx);
        const uint32_t cpl = (uint32_t)(g.py * ngx);
        const uint32_t cA = ((k * cpl + cj) << 1) | (uint32_t)iparA, cB = ((kb * cpl + cj) << 1) | (uint32_t)iparB;
        uint32_t k0 = nk.k0, k1 = nk.k1;  // the key schedule per step (opaque_sgpr)
        opaque_sgpr(k0);
        opaque_sgpr(k1);
        if (doA) normals4(cA, (uint32_t)chain, nk.tag, sample, k0, k1, xiA);
        if (doB) normals4(cB, (uint32_t)chain, nk.tag, sample, k0, k1, xiB);
        if constexpr (CB > 1) {
          if (doA) normals4(cA, (uint32_t)(chain + 1), nk.tag, sample, k0, k1, xiA1);
          if (doB) normals4(cB, (uint32_t)(chain + 1), nk.tag, sample, k0, k1, xiB1);
        }
      }
      if (next) {
        pro.apply(xp, k + 1);
        if constexpr (CB > 1) pro1.apply(xp1, k + 1);
      }
      if (active) LdsRow<T>::store(A0 + (k & 1) * asz + aoff, gx, ngx, xc);  // plane k before the first colour
      if constexpr (CB > 1) {
        if (active) LdsRow<T>::store(A0 + csz + (k & 1) * asz + aoff, gx, ngx, xc1);
      }
      __syncthreads();
#define MGMC_RB3_UPDATE(XMM, XM, XC, XP, XIA, XIB, CBASE, AC)                                                      \
    {                                                                                                                \
      const T* A = (AC) + (k & 1) * asz;                                                                             \
      const T* Bm = (AC) + 2 * asz + (kb & 1) * bsz; /* plane k-1 after the first colour (written last step) */     \
      T* Bk = (AC) + 2 * asz + (k & 1) * bsz;                                                                        \
      /* first colour on plane k (owned rows and halo ring) */                                                       \
      if (doA) {                                                                                                     \
        const T* Arow = A + aoff;                                                                                    \
        T ym[8], yp[8];                                                                                              \
        LdsRow<T>::load(Arow - g.px, gx, ngx, ym);                                                                   \
        LdsRow<T>::load(Arow + g.px, gx, ngx, yp);                                                                   \
        const T left = gx > 0 ? LdsRow<T>::left(Arow, gx, ngx) : (T)0;                                               \
        const T right = gx < ngx - 1 ? LdsRow<T>::right(Arow, gx, ngx) : (T)0;                                       \
        const T* eg = NOISE == NOISE_EXTERNAL ? ext + CBASE + (long long)k * pl : nullptr;                           \
        if (iparA == 0)                                                                                              \
          star_update4<0, T, NOISE>(XC, left, right, ym, yp, XM, XP, fc, XIA, eg, 8 * gx, g.n0, cx, cy, cz, D, om_d, \
                                    ns);                                                                             \
        else                                                                                                         \
          star_update4<1, T, NOISE>(XC, left, right, ym, yp, XM, XP, fc, XIA, eg, 8 * gx, g.n0, cx, cy, cz, D, om_d, \
                                    ns);                                                                             \
      }                                                                                                              \
      /* second colour on plane k-1 (owned rows, owned planes) */                                                    \
      if (doB) {                                                                                                     \
        const T* Brow = Bm + boff;                                                                                   \
        T ym[8], yp[8];                                                                                              \
        LdsRow<T>::load(Brow - g.px, gx, ngx, ym);                                                                   \
        LdsRow<T>::load(Brow + g.px, gx, ngx, yp);                                                                   \
        const T left = gx > 0 ? LdsRow<T>::left(Brow, gx, ngx) : (T)0;                                               \
        const T right = gx < ngx - 1 ? LdsRow<T>::right(Brow, gx, ngx) : (T)0;                                       \
        const T* eg = NOISE == NOISE_EXTERNAL ? ext + CBASE + (long long)kb * pl : nullptr;                          \
        if (iparB == 0)                                                                                              \
          star_update4<0, T, NOISE>(XM, left, right, ym, yp, XMM, XC, fm, XIB, eg, 8 * gx, g.n0, cx, cy, cz, D,      \
                                    om_d, ns);                                                                       \
        else                                                                                                         \
          star_update4<1, T, NOISE>(XM, left, right, ym, yp, XMM, XC, fm, XIB, eg, 8 * gx, g.n0, cx, cy, cz, D,      \
                                    om_d, ns);                                                                       \
      }                                                                                                              \
      if (storeB) store8<T>(xout + CBASE + (long long)kb * pl, XM);                                                  \
      /* plane k after the first colour, for the second colour at the next step */                                   \
      if (active && r >= 1 && r <= TY + 2) LdsRow<T>::store(Bk + boff, gx, ngx, XC);                                 \
      _Pragma("unroll") for (int q = 0; q < 8; ++q) {                                                                \
        XMM[q] = XM[q];                                                                                              \
        XM[q] = XC[q];                                                                                               \
        XC[q] = XP[q];                                                                                               \
      }                                                                                                              \
    }
      MGMC_RB3_UPDATE(xmm, xm, xc, xp, xiA, xiB, cbase, A0)
      if constexpr (CB > 1) {
        MGMC_RB3_UPDATE(xmm1, xm1, xc1, xp1, xiA1, xiB1, cbase1, A0 + csz)
      }
#undef MGMC_RB3_UPDATE
#pragma unroll
      for (int q = 0; q < 8; ++q) fm[q] = fc[q];
    }
  };

  // nseg > 0: workgroup (chain group * nseg + seg) * ntile + tile marches the
  // planes [1 + seg * SZ, 1 + (seg + 1) * SZ). nseg = 0: the ncol columns
  // (chain group * ntile + tile) times the n2 - 1 interior planes, in that
  // order, are split into equal runs, one per workgroup of the grid; a run
  // that crosses into the next column is marched as two segments. The values
  // written do not depend on where the segments end.
  // (the host keeps ncol * (n2 - 1) below 2^31; the run bounds are uniform)
  const int P = g.n2 - 1;
  int a, e;
  if (nseg > 0) {
    const int b = (int)bk.x / ntile;
    a = (b / nseg * ntile + (int)bk.x % ntile) * P + (b % nseg) * SZ;
    e = a + SZ < (a / P + 1) * P ? a + SZ : (a / P + 1) * P;
  } else {
    // equal runs: the first total % grid workgroups take one unit more
    const int total = ncol * P, q = total / (int)bk.gx, rem = total % (int)bk.gx, w = (int)bk.x;
    a = w * q + (w < rem ? w : rem);
    e = a + q + (w < rem ? 1 : 0);
  }
  a = __builtin_amdgcn_readfirstlane(a);
  e = __builtin_amdgcn_readfirstlane(e);
  for (bool firstseg = true; a < e; firstseg = false) {
    const int col = a / P;
    const int p1 = e - col * P < P ? e - col * P : P;
    if (!firstseg) __syncthreads();  // the previous segment's last LDS reads
    march(col % ntile, col / ntile * CB, 1 + a - col * P, 1 + p1);
    a = col * P + p1;
  }
}
template <typename T, int NOISE, int CB>
__global__ void __launch_bounds__((rb3_max_threads<T, CB>())) k_sweep_rb3(const T* __restrict__ x, T* __restrict__ xout,
                                                    const T* __restrict__ f, const T* __restrict__ ext, DevGeom g,
                                                    Stencil st, int first, int TY, int SZ, int ntile, int nseg,
                                                    int ncol, NoiseKey nk, const T* __restrict__ xcoarse, DevGeom gc,
                                                    T alpha) {
  b_sweep_rb3<T, NOISE, CB>(Blk::hw(), x, xout, f, ext, g, st, first, TY, SZ, ntile, nseg, ncol, nk, xcoarse, gc,
                            alpha);
}
template <typename T, int NOISE, int CB>
constexpr Batchable<b_sweep_rb3<T, NOISE, CB>, (rb3_max_threads<T, CB>()), true, k_sweep_rb3<T, NOISE, CB>> kb_sweep_rb3{};

// ---------------------------------------------------------------------------
// Half of an 8-colour SOR/Gibbs sweep of a 3D box-stencil (27-point) level:
// every interior plane of z-parity `kpar` gets its four colours (ipar, jpar)
// in sweep order. Planes of equal parity do not couple (the stencil reaches
// k +- 1 only), so one workgroup per (plane, chain) reproduces the colour-
// phase launches of k_sweep for colours 4 kpar .. 4 kpar + 3 (forward) or the
// reverse (backward). Plane k lives in LDS; planes k +- 1 are read-only here.
// One thread per aligned 8-vertex group of a row; the first `hth` threads
// (a multiple of 64) hold the odd rows and the rest the even rows, so that
// the waves of the row parity a colour phase does not touch skip it whole.
// ---------------------------------------------------------------------------
template <typename T, int NOISE, bool STAGE3>
__device__ __forceinline__ void b_sweep_box_planes(const Blk bk, T* __restrict__ x, const T* __restrict__ f,
                                                           const T* __restrict__ ext, DevGeom g, StencilT<T> st,
                                                           const T* __restrict__ coef, int kpar, int forward,
                                                           int nplanes, int hth, int xzero, NoiseKey nk) {
  MGMC_DYNAMIC_SHARED(char, lds_raw);
  // STAGE3: planes k-1, k, k+1 in LDS (py x px each, LdsRow chunks); else
  // plane k only (plain rows: planes k +- 1 are read from global memory)
  constexpr bool CH = STAGE3;
  const size_t psz = (size_t)g.py * g.px;
  T* P = reinterpret_cast<T*>(lds_raw) + (STAGE3 ? psz : 0);
  const int ngx = g.px >> 3;
  const int tid = threadIdx.x;
  const bool odd = tid < hth;
  const int t2 = odd ? tid : tid - hth;
  const int gx = t2 % ngx;
  const int j = 2 * (t2 / ngx) + (odd ? 1 : 2);
  const bool active = j <= g.n1 - 1;
  // workgroups are dispatched to the eight XCDs round-robin: deal out the
  // (plane, chain) items so that an XCD runs consecutive planes of a chain,
  // whose shared neighbour planes then meet in its L2 (FP32 cycle at 128^3
  // x 64 chains: -1.2 % against the plain order)
  int w = (int)bk.x;
  const int nb = (int)bk.gx;
  if ((nb & 7) == 0) w = (w & 7) * (nb >> 3) + (w >> 3);
  const int pidx = w % nplanes;
  const int chain = w / nplanes;
  const int k = kpar ? 2 * pidx + 1 : 2 * pidx + 2;
  const long long base = (long long)chain * g.cs + ((long long)k * g.py + j) * g.px + 8 * gx;
  const uint32_t sample = *nk.sample;
  const int jpar = j & 1;
  // stage plane k (and k +- 1), boundary rows 0 and n1 included, in LDS: the
  // loads are issued first and the normals of the thread's two colour phases
  // are drawn while they are in flight
  // xzero (STAGE3 only): 1 = every plane of x is zero, 2 = the planes of
  // parity kpar are zero; zero planes are not read
  const int ngrp = g.py * ngx;
  const bool one = ngrp <= (int)blockDim.x;  // one group per thread and plane
  T v[STAGE3 ? 3 : 1][8];
  const int lr = tid / ngx, lg = tid % ngx;
  if (one && tid < ngrp)
    for (int d = 0; d < (STAGE3 ? 3 : 1); ++d) {
      const bool own = !STAGE3 || d == 1;
      if (xzero == 1 || (xzero == 2 && own)) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[d][q] = (T)0;
      } else {
        load8<T>(x + (long long)chain * g.cs + ((long long)(k + d - (STAGE3 ? 1 : 0)) * g.py + lr) * g.px + 8 * lg,
                 v[d]);
      }
    }
  T fb[8];
  if (active) load8<T>(f + base + (long long)chain * (g.fcs - g.cs), fb);
  float xi[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  if (NOISE == NOISE_PHILOX && active && 8 * gx <= g.n0 - 1)
    for (int ipar = 0; ipar < 2; ++ipar) {
      uint32_t c0 = (uint32_t)((((long long)k * g.py + j) * ngx + gx) << 1) | (uint32_t)ipar;
      normals4(c0, (uint32_t)chain, nk.tag, sample, nk.k0, nk.k1, xi[ipar]);
    }
  if (one) {
    if (tid < ngrp)
      for (int d = 0; d < (STAGE3 ? 3 : 1); ++d) {
        T* q = P + (d - (STAGE3 ? 1 : 0)) * (long long)psz + (size_t)lr * g.px;
        if (CH) LdsRow<T>::store(q, lg, ngx, v[d]);
        else store8<T>(q + 8 * lg, v[d]);
      }
  } else {
    for (int dz = STAGE3 ? -1 : 0; dz <= (STAGE3 ? 1 : 0); ++dz)
      for (int e = tid; e < ngrp; e += blockDim.x) {
        const int rr = e / ngx, gg = e % ngx;
        T w[8];
        if (xzero == 1 || (xzero == 2 && dz == 0)) {
#pragma unroll
          for (int q = 0; q < 8; ++q) w[q] = (T)0;
        } else {
          load8<T>(x + (long long)chain * g.cs + ((long long)(k + dz) * g.py + rr) * g.px + 8 * gg, w);
        }
        T* q = P + dz * (long long)psz + (size_t)rr * g.px;
        if (CH) LdsRow<T>::store(q, gg, ngx, w);
        else store8<T>(q + 8 * gg, w);
      }
  }
  const T* eg = NOISE == NOISE_EXTERNAL ? ext + base : nullptr;
  for (int ph = 0; ph < 4; ++ph) {
    const int q = forward ? ph : 3 - ph;  // in-plane colour: ipar = q & 1, jpar = q >> 1
    __syncthreads();
    if (!active || jpar != (q >> 1) || 8 * gx > g.n0 - 1) continue;
    const int ipar = q & 1;
    T* row = P + (size_t)j * g.px + (CH ? 0 : 8 * gx);
    T xc[10];
    if (CH) {
      LdsRow<T>::load(row, gx, ngx, xc + 1);
      xc[0] = gx > 0 ? LdsRow<T>::left(row, gx, ngx) : (T)0;
      xc[9] = gx < ngx - 1 ? LdsRow<T>::right(row, gx, ngx) : (T)0;
    } else {
      load8<T>(row, xc + 1);
      xc[0] = gx > 0 ? row[-1] : (T)0;
      xc[9] = gx < ngx - 1 ? row[8] : (T)0;
    }
    const T* zm = STAGE3 ? row - psz : x + base - (long long)g.px * g.py;
    const T* zp = STAGE3 ? row + psz : x + base + (long long)g.px * g.py;
    if (ipar == 0)
      box_plane_update4<0, T, NOISE, CH>(xc, row, zm, zp, g, st, coef, fb, xi[0], eg, 8 * gx, gx, ngx);
    else
      box_plane_update4<1, T, NOISE, CH>(xc, row, zm, zp, g, st, coef, fb, xi[1], eg, 8 * gx, gx, ngx);
    if (CH) LdsRow<T>::store(row, gx, ngx, xc + 1);
    else store8<T>(row, xc + 1);
  }
  __syncthreads();
  if (active) {
    T v[8];
    if (CH) LdsRow<T>::load(P + (size_t)j * g.px, gx, ngx, v);
    else load8<T>(P + (size_t)j * g.px + 8 * gx, v);
    store8<T>(x + base, v);
  }
}
template <typename T, int NOISE, bool STAGE3>
__global__ void __launch_bounds__(1024) k_sweep_box_planes(T* __restrict__ x, const T* __restrict__ f,
                                                           const T* __restrict__ ext, DevGeom g, StencilT<T> st,
                                                           const T* __restrict__ coef, int kpar, int forward,
                                                           int nplanes, int hth, int xzero, NoiseKey nk) {
  b_sweep_box_planes<T, NOISE, STAGE3>(Blk::hw(), x, f, ext, g, st, coef, kpar, forward, nplanes, hth, xzero, nk);
}
template <typename T, int NOISE, bool STAGE3>
constexpr Batchable<b_sweep_box_planes<T, NOISE, STAGE3>, 1024, true, k_sweep_box_planes<T, NOISE, STAGE3>> kb_sweep_box_planes{};

// ---------------------------------------------------------------------------
// One whole 8-colour SOR/Gibbs direction of a 3D box-stencil (27-point) level
// in a single pass: a workgroup owns a z-segment [ks, ke) of one chain and
// marches over the planes of the parity swept first (P1: even planes going
// forward, odd going backward). At P1 plane k it applies that plane's four
// colours (neighbours: the old planes k +- 1), then the four colours of plane
// k-1 (neighbours: the new planes k-2 and k). Four planes live in an LDS ring
// (k-2 .. k+1); planes k+2, k+3 are loaded into registers while the eight
// colour phases run. P1 planes just outside the segment are recomputed
// (old inputs and vertex-keyed noise: the owner gets the same values) and
// the sweep writes to a second buffer, so every workgroup reads old values
// only. Result = the eight colour-phase launches of k_sweep. Rows are
// parity-sorted as in k_sweep_box_planes; one thread per 8-vertex group.
// ---------------------------------------------------------------------------
// normals of the two colour phases of the thread's row on plane kz
template <int NOISE>
__device__ __forceinline__ void box_row_normals(int kz, const DevGeom& g, int chain, int j, int gx, int ngx,
                                                bool active, uint32_t sample, const NoiseKey& nk, float (&xi)[2][4]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) xi[a][b] = 0.f;
  if (NOISE != NOISE_PHILOX || !active) return;
  for (int ipar = 0; ipar < 2; ++ipar) {
    uint32_t c0 = (uint32_t)((((long long)kz * g.py + j) * ngx + gx) << 1) | (uint32_t)ipar;
    normals4(c0, (uint32_t)chain, nk.tag, sample, nk.k0, nk.k1, xi[ipar]);
  }
}

template <typename T, int NOISE>
__device__ __forceinline__ void box_march_plane(T* __restrict__ R, size_t psz, int kz, const DevGeom& g,
                                                const StencilT<T>& st, const T* __restrict__ coef, const T* fb,
                                                const T* __restrict__ ext, long long cb, int j, int gx, int ngx,
                                                bool active, int forward, const float (&xi)[2][4]) {
  T* P = R + (size_t)(kz & 3) * psz;
  const T* zm = R + (size_t)((kz - 1) & 3) * psz + (size_t)j * g.px + 8 * gx;
  const T* zp = R + (size_t)((kz + 1) & 3) * psz + (size_t)j * g.px + 8 * gx;
  const int jpar = j & 1;
  const T* eg = NOISE == NOISE_EXTERNAL ? ext + cb + ((long long)kz * g.py + j) * g.px + 8 * gx : nullptr;
  for (int ph = 0; ph < 4; ++ph) {
    const int q = forward ? ph : 3 - ph;  // in-plane colour: ipar = q & 1, jpar = q >> 1
    __syncthreads();
    if (!active || jpar != (q >> 1)) continue;
    const int ipar = q & 1;
    T* row = P + (size_t)j * g.px + 8 * gx;
    T xc[10];
    load8<T>(row, xc + 1);
    xc[0] = gx > 0 ? row[-1] : (T)0;
    xc[9] = gx < ngx - 1 ? row[8] : (T)0;
    if (ipar == 0)
      box_plane_update4<0, T, NOISE, false>(xc, row, zm, zp, g, st, coef, fb, xi[0], eg, 8 * gx, gx, ngx);
    else
      box_plane_update4<1, T, NOISE, false>(xc, row, zm, zp, g, st, coef, fb, xi[1], eg, 8 * gx, gx, ngx);
    store8<T>(row, xc + 1);
  }
}

template <typename T, int NOISE>
__global__ void __launch_bounds__(1024) k_sweep_box_march(const T* __restrict__ x, T* __restrict__ xout,
                                                          const T* __restrict__ f, const T* __restrict__ ext,
                                                          DevGeom g, StencilT<T> st, const T* __restrict__ coef,
                                                          int forward, int SZ, int nseg, int hth, NoiseKey nk) {
  MGMC_DYNAMIC_SHARED(char, lds_raw);
  T* R = reinterpret_cast<T*>(lds_raw);  // four plane slots, plane p in slot p & 3
  const size_t psz = (size_t)g.py * g.px;
  const int ngx = g.px >> 3;
  const int tid = threadIdx.x;
  const bool odd = tid < hth;
  const int t2 = odd ? tid : tid - hth;
  const int gx = t2 % ngx;
  const int j = 2 * (t2 / ngx) + (odd ? 1 : 2);
  const bool rowok = j <= g.n1 - 1;
  const bool active = rowok && 8 * gx <= g.n0 - 1;
  const int seg = blockIdx.x % nseg;
  const int chain = blockIdx.x / nseg;
  const int ks = 1 + seg * SZ;
  const int ke = ks + SZ < g.n2 ? ks + SZ : g.n2;  // owned planes [ks, ke)
  const int p1 = forward ? 0 : 1;                  // parity of the planes swept first
  const long long cb = (long long)chain * g.cs;
  const long long fcb = (long long)chain * g.fcs;
  const uint32_t sample = *nk.sample;
  // the thread's group of each plane it loads (one 8-vertex group: py * ngx <= threads)
  const int lr = tid / ngx, lg = tid % ngx;
  const bool loader = lr < g.py;
  const long long loff = (long long)lr * g.px + 8 * lg;
  int kA = ks - 1;
  if ((kA & 1) != p1) ++kA;  // first P1 plane: ks - 1 or ks
  int kB = ke;
  if ((kB & 1) != p1) --kB;  // last P1 plane: ke or ke - 1
  // initial ring: old planes kA - 1, kA, kA + 1
  for (int d = -1; d <= 1; ++d) {
    const int p = kA + d;
    if (loader) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (T)0;
      if (p >= 0 && p <= g.n2) load8<T>(x + cb + (long long)p * psz + loff, v);
      store8<T>(R + (size_t)(p & 3) * psz + loff, v);
    }
  }
  for (int k = kA; k <= kB; k += 2) {
    // planes k+2, k+3 in flight during the colour phases of this step
    T n2v[8], n3v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) n2v[q] = n3v[q] = (T)0;
    const bool more = k + 2 <= kB;
    if (loader && more) {
      if (k + 2 <= g.n2) load8<T>(x + cb + (long long)(k + 2) * psz + loff, n2v);
      if (k + 3 <= g.n2) load8<T>(x + cb + (long long)(k + 3) * psz + loff, n3v);
    }
    const bool up1 = k >= 1 && k <= g.n2 - 1;
    const int kb = k - 1;
    const bool up2 = kb >= ks && kb < ke && kb >= 1 && kb <= g.n2 - 1;
    T f1[8], f2[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) f1[q] = f2[q] = (T)0;
    if (active && up1) load8<T>(f + fcb + ((long long)k * g.py + j) * g.px + 8 * gx, f1);
    if (active && up2) load8<T>(f + fcb + ((long long)kb * g.py + j) * g.px + 8 * gx, f2);
    // the normals of both planes while the loads are in flight
    float xi1[2][4], xi2[2][4];
    box_row_normals<NOISE>(k, g, chain, j, gx, ngx, active && up1, sample, nk, xi1);
    box_row_normals<NOISE>(kb, g, chain, j, gx, ngx, active && up2, sample, nk, xi2);
    if (up1) box_march_plane<T, NOISE>(R, psz, k, g, st, coef, f1, ext, cb, j, gx, ngx, active, forward, xi1);
    if (up2) box_march_plane<T, NOISE>(R, psz, kb, g, st, coef, f2, ext, cb, j, gx, ngx, active, forward, xi2);
    __syncthreads();
    // owned planes of this step to the output buffer (whole rows, padding included)
    if (rowok) {
      const long long roff = (long long)j * g.px + 8 * gx;
      if (k >= ks && k < ke) {
        T v[8];
        load8<T>(R + (size_t)(k & 3) * psz + roff, v);
        store8<T>(xout + cb + (long long)k * psz + roff, v);
      }
      if (kb >= ks && kb < ke) {
        T v[8];
        load8<T>(R + (size_t)(kb & 3) * psz + roff, v);
        store8<T>(xout + cb + (long long)kb * psz + roff, v);
      }
    }
    __syncthreads();
    // planes k-2 and k-1 are done: their slots take k+2 and k+3
    if (loader && more) {
      store8<T>(R + (size_t)((k + 2) & 3) * psz + loff, n2v);
      store8<T>(R + (size_t)((k + 3) & 3) * psz + loff, n3v);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused residual and restriction on a 3D star-stencil level:
//   fc = R (f - A x)   (multigridmc_sampler.cc:125-128 without the low-rank
//   term, which is added by the low-rank kernels).
// A workgroup owns TYc coarse rows (fine rows 2 J0 - 1 .. 2 J0 + 2 TYc - 1
// plus one more x row on each side) and marches over the fine planes of its
// coarse z-segment [Ks, Ke): r(p) is formed from plane p in LDS and planes
// p +- 1 in registers, restricted in x (left neighbour from LDS), in y (LDS)
// and accumulated in z; coarse plane K is stored when fine plane 2K + 1 is
// done. One thread per fine 8-vertex group and fine row. Segments as in
// k_sweep_rb3: nseg > 0 gives z-segments of SZc coarse planes, nseg = 0
// splits the ncol columns (chain * ntile + tile) x coarse planes evenly over
// the grid.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void b_residual_restrict_star3(const Blk bk, const T* __restrict__ x, const T* __restrict__ f,
                                                                  T* __restrict__ fc, DevGeom g, DevGeom gc,
                                                                  Stencil st, int TYc, int SZc, int ntile, int nseg,
                                                                  int ncol) {
  MGMC_DYNAMIC_SHARED(char, lds_raw);
  const int nrx = 2 * TYc + 3, nrr = 2 * TYc + 1;
  T* LX = reinterpret_cast<T*>(lds_raw);    // nrx rows x px: x plane p
  T* LR = LX + (size_t)nrx * g.px;          // nrr rows x px: r plane p
  T* LT = LR + (size_t)nrr * g.px;          // nrr rows x (px / 2): x-restricted r
  const int ngx = g.px >> 3;
  const int tid = threadIdx.x;
  const int rt = tid / ngx, gx = tid % ngx;
  const bool active = rt < nrx;
  const bool rrow = active && rt >= 1 && rt <= nrr;                 // rows whose residual enters the tile
  const long long pl = (long long)g.px * g.py;
  const T cx = (T)st.c[12], cy = (T)st.c[10], cz = (T)st.c[4], D = (T)st.diag;
  const T half = (T)0.5;

  // the march of chain `chain` over tile `tile`, coarse planes [Ks, Ke)
  auto march = [&](const int tile, const int chain, const int Ks, const int Ke) {
    const int J0 = 1 + tile * TYc;
    const int j = 2 * J0 - 2 + rt;
    const bool has_mem = active && j >= 0 && j <= g.n1;
    const bool rvalid = rrow && j >= 1 && j <= g.n1 - 1;
    const int J = j >> 1;
    const bool crow = rrow && !(j & 1) && J >= J0 && J < J0 + TYc && J <= gc.n1 - 1;
    const long long cbase = (long long)chain * g.cs + (long long)j * g.px + 8 * gx;
    const long long fbase = (long long)chain * g.fcs + (long long)j * g.px + 8 * gx;

    T xm[8], xc[8], xp[8], xq[8], fb[8], fn[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) xm[q] = xc[q] = xp[q] = xq[q] = fb[q] = fn[q] = (T)0;
    const int p0 = 2 * Ks - 1, p1 = 2 * Ke - 1;  // p1 + 1 <= n2
    if (has_mem) {
      load8<T>(x + cbase + (long long)(p0 - 1) * pl, xm);
      load8<T>(x + cbase + (long long)p0 * pl, xc);
      load8<T>(x + cbase + (long long)(p0 + 1) * pl, xp);
    }
    if (rvalid) load8<T>(f + fbase + (long long)p0 * pl, fb);
    T acc[4] = {(T)0, (T)0, (T)0, (T)0};
    T* LXrow = LX + (size_t)rt * g.px;  // LdsRow chunks
    T* LRrow = LR + (size_t)(rt - 1) * g.px;
    T* LTrow = LT + (size_t)(rt - 1) * (g.px >> 1) + 4 * gx;

    for (int p = p0; p <= p1; ++p) {
      // plane p + 2 of x and plane p + 1 of f are in flight while plane p is
      // processed (x two planes ahead: plane p + 1 is read in this step)
      if (has_mem && p + 2 <= g.n2) load8<T>(x + cbase + (long long)(p + 2) * pl, xq);
      if (rvalid && p < p1) load8<T>(f + fbase + (long long)(p + 1) * pl, fn);
      if (active) LdsRow<T>::store(LXrow, gx, ngx, xc);
      __syncthreads();
      T r[8];
      if (rrow) {
        T ym[8], yp[8];
        LdsRow<T>::load(LXrow - g.px, gx, ngx, ym);
        LdsRow<T>::load(LXrow + g.px, gx, ngx, yp);
        const T left = gx > 0 ? LdsRow<T>::left(LXrow, gx, ngx) : (T)0;
        const T right = gx < ngx - 1 ? LdsRow<T>::right(LXrow, gx, ngx) : (T)0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const T xl = q > 0 ? xc[q - 1] : left;
          const T xr = q < 7 ? xc[q + 1] : right;
          T a = cy * ym[q] + cx * xl + D * xc[q] + cx * xr + cy * yp[q];
          a = cz * xm[q] + a + cz * xp[q];
          const int i = 8 * gx + q;
          r[q] = (rvalid && i >= 1 && i <= g.n0 - 1) ? fb[q] - a : (T)0;
        }
        LdsRow<T>::store(LRrow, gx, ngx, r);
      }
      __syncthreads();
      if (rrow) {
        const T rl = gx > 0 ? LdsRow<T>::left(LRrow, gx, ngx) : (T)0;
        T t[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) t[a] = half * (a > 0 ? r[2 * a - 1] : rl) + r[2 * a] + half * r[2 * a + 1];
#pragma unroll
        for (int a = 0; a < 4; ++a) LTrow[a] = t[a];
      }
      __syncthreads();
      if (crow) {
        const int hp = g.px >> 1;
        T u[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) u[a] = half * LTrow[a - hp] + LTrow[a] + half * LTrow[a + hp];
        if (p & 1) {
          // fine plane 2K+1 completes coarse plane K = (p-1)/2 and starts (p+1)/2
          const int K = (p - 1) >> 1;
          if (K >= Ks) {
            T* out = fc + (long long)chain * gc.cs + ((long long)K * gc.py + J) * gc.px + 4 * gx;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              const int I = 4 * gx + a;
              if (I >= 1 && I <= gc.n0 - 1) out[a] = acc[a] + half * u[a];
            }
          }
#pragma unroll
          for (int a = 0; a < 4; ++a) acc[a] = half * u[a];
        } else {
#pragma unroll
          for (int a = 0; a < 4; ++a) acc[a] += u[a];
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        xm[q] = xc[q];
        xc[q] = xp[q];
        xp[q] = xq[q];
        fb[q] = fn[q];
      }
    }
  };

  // (the host keeps ncol * (gc.n2 - 1) below 2^31; the run bounds are uniform)
  const int P = gc.n2 - 1;
  int a, e;
  if (nseg > 0) {
    const int b = (int)bk.x / ntile;
    a = (b / nseg * ntile + (int)bk.x % ntile) * P + (b % nseg) * SZc;
    e = a + SZc < (a / P + 1) * P ? a + SZc : (a / P + 1) * P;
  } else {
    // equal runs: the first total % grid workgroups take one unit more
    const int total = ncol * P, q = total / (int)bk.gx, rem = total % (int)bk.gx, w = (int)bk.x;
    a = w * q + (w < rem ? w : rem);
    e = a + q + (w < rem ? 1 : 0);
  }
  a = __builtin_amdgcn_readfirstlane(a);
  e = __builtin_amdgcn_readfirstlane(e);
  for (bool firstseg = true; a < e; firstseg = false) {
    const int col = a / P;
    const int p1 = e - col * P < P ? e - col * P : P;
    if (!firstseg) __syncthreads();  // the previous segment's last LDS reads
    march(col % ntile, col / ntile, 1 + a - col * P, 1 + p1);
    a = col * P + p1;
  }
}
template <typename T>
__global__ void __launch_bounds__(rb3_max_threads<T>()) k_residual_restrict_star3(const T* __restrict__ x, const T* __restrict__ f,
                                                                  T* __restrict__ fc, DevGeom g, DevGeom gc,
                                                                  Stencil st, int TYc, int SZc, int ntile, int nseg,
                                                                  int ncol) {
  b_residual_restrict_star3<T>(Blk::hw(), x, f, fc, g, gc, st, TYc, SZc, ntile, nseg, ncol);
}
template <typename T>
constexpr Batchable<b_residual_restrict_star3<T>, rb3_max_threads<T>(), true, k_residual_restrict_star3<T>> kb_residual_restrict_star3{};

// obs[row][c][p] = x_c[probe_p], row taken from the device run counter
// obs[row][c][p] = sum_e w_e x_c[idx_e] over the entries of observable p
// (a probe vertex is one entry of weight 1; a ball average is a measurement
// vector), row taken from the device run counter. One wave per (p, c).
template <typename T>
__device__ __forceinline__ void b_observe(const Blk bk, const T* __restrict__ x, long long cs,
                                                const int32_t* __restrict__ optr,
                                                const long long* __restrict__ oidx, const double* __restrict__ ow,
                                                int P, int K, double* __restrict__ obs,
                                                const uint32_t* __restrict__ ctr) {
  const int p = bk.x, c = bk.y;
  const T* xc = x + (long long)c * cs;
  double s = 0.0;
  for (int e = optr[p] + (int)threadIdx.x; e < optr[p + 1]; e += 64) s += ow[e] * (double)xc[oidx[e]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (threadIdx.x == 0) obs[((long long)ctr[1] * K + c) * P + p] = s;
}
template <typename T>
__global__ void __launch_bounds__(64) k_observe(const T* __restrict__ x, long long cs,
                                                const int32_t* __restrict__ optr,
                                                const long long* __restrict__ oidx, const double* __restrict__ ow,
                                                int P, int K, double* __restrict__ obs,
                                                const uint32_t* __restrict__ ctr) {
  b_observe<T>(Blk::hw(), x, cs, optr, oidx, ow, P, K, obs, ctr);
}
template <typename T>
constexpr Batchable<b_observe<T>, 64, false, k_observe<T>> kb_observe{};

// m1[i] += sum_c x_c[i], m2[i] += sum_c x_c[i]^2 over the padded level-0 box
// (pointwise moments for the posterior mean and variance fields)
template <typename T>
__global__ void __launch_bounds__(256) k_moments(const T* __restrict__ x, long long cs, int K, long long vol,
                                                 double* __restrict__ m1, double* __restrict__ m2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vol) return;
  double s1 = 0.0, s2 = 0.0;
  for (int c = 0; c < K; ++c) {
    const double v = (double)x[c * cs + i];
    s1 += v;
    s2 += v * v;
  }
  m1[i] += s1;
  m2[i] += s2;
}

// advance the sample counter and the observable row after a cycle
__device__ __forceinline__ void b_advance(const Blk bk, uint32_t* ctr) {
  if (threadIdx.x == 0 && bk.x == 0) {
    ctr[0] += 1u;
    ctr[1] += 1u;
  }
}
__global__ void __launch_bounds__(1024) k_advance(uint32_t* ctr) {
  b_advance(Blk::hw(), ctr);
}
constexpr Batchable<b_advance, 1024, false, k_advance> kb_advance{};

// The normal the sweeps draw at every fine vertex (reference order)
__global__ void k_normals(double* __restrict__ out, DevGeom g, int K, NoiseKey nk) {
  long long per = (long long)(g.n0 - 1) * (g.n1 - 1) * (g.dim == 3 ? g.n2 - 1 : 1);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * K) return;
  int c = (int)(t / per);
  long long e = t % per;
  int i = (int)(e % (g.n0 - 1)) + 1;
  e /= (g.n0 - 1);
  int j = (int)(e % (g.n1 - 1)) + 1;
  int k = g.dim == 3 ? (int)(e / (g.n1 - 1)) + 1 : 0;
  int ngx = g.px >> 3;
  uint32_t c0 = (uint32_t)((((long long)k * g.py + j) * ngx + (i >> 3)) << 1) | (uint32_t)(i & 1);
  float xi[4];
  normals4(c0, (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
  out[t] = (double)xi[(i & 7) >> 1];
}

// FP64 <-> T copies and updates used by the loop solver
template <typename T>
__global__ void k_axpy_pad(double* __restrict__ y, const T* __restrict__ x, double a, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] += a * (double)x[t];
}
template <typename T>
__global__ void k_convert(T* __restrict__ y, const double* __restrict__ x, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (T)x[t];
}
// per-chain partial sums of squares: chain blockIdx.y, out[chain][block]
__global__ void k_sumsq(const double* __restrict__ x, long long n, double* __restrict__ out) {
  double s = 0.0;
  x += (long long)blockIdx.y * n;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    s += x[t] * x[t];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[4];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[(long long)blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Contraction of adjoint and forward fields for parameter gradients: for each
// chain c, out[c][0] = sum_v lam_c[v] mu_c[v] and out[c][1] = sum_v lam_c[v]
// (S mu_c)[v], S a constant 3^d stencil applied with Dirichlet boundaries (the
// derivative of the operator with respect to a stencil parameter). Padded
// FP64 fields; one 8-vertex group per thread, grid (row blocks, chains),
// per-block partials out_part[c][block][2] summed on the host.
__global__ void __launch_bounds__(256) k_contract(const double* __restrict__ lam, const double* __restrict__ mu,
                                                  DevGeom g, StencilT<double> S, double* __restrict__ out_part) {
  const int ngx = g.px >> 3;
  const int nrows = (g.n1 - 1) * (g.dim == 3 ? g.n2 - 1 : 1);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  double s0 = 0.0, s1 = 0.0;
  if (t < ngx * nrows) {
    const int gx = t % ngx, row = t / ngx;
    const int j = row % (g.n1 - 1) + 1;
    const int k = g.dim == 3 ? row / (g.n1 - 1) + 1 : 0;
    const long long pl = (long long)g.px * g.py;
    const long long base = (long long)c * g.cs + ((long long)k * g.py + j) * g.px + 8 * gx;
    const int zr = g.dim == 3 ? 1 : 0;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int dz = -zr; dz <= zr; ++dz)
      for (int dy = -1; dy <= 1; ++dy) {
        const double* p = mu + base + dz * pl + dy * g.px;
        double w[10];
        load8<double>(p, w + 1);
        w[0] = p[-1];
        w[9] = p[8];
        const int cb = (g.dim == 3 ? (dz + 1) * 9 : 0) + (dy + 1) * 3;
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] += S.c[cb] * w[q] + S.c[cb + 1] * w[q + 1] + S.c[cb + 2] * w[q + 2];
      }
    double lv[8], mv[8];
    load8<double>(lam + base, lv);
    load8<double>(mu + base, mv);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = 8 * gx + q;
      if (i >= 1 && i <= g.n0 - 1) {
        s0 += lv[q] * mv[q];
        s1 += lv[q] * acc[q];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  __shared__ double red[2][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = s0;
    red[1][wid] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = out_part + ((long long)c * gridDim.x + blockIdx.x) * 2;
    o[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    o[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// Reference vertex index of a padded-layout index (interior vertices only;
// src/lattice/lattice3d.hh:132-141).
__device__ __forceinline__ long long padded_to_ref(long long p, const DevGeom& g) {
  const int i = (int)(p % g.px);
  const long long q = p / g.px;
  const int j = (int)(q % g.py);
  const int k = (int)(q / g.py);
  const long long row = g.dim == 3 ? (long long)(k - 1) * (g.n1 - 1) + (j - 1) : (long long)(j - 1);
  return row * (g.n0 - 1) + (i - 1);
}

// z[c][k] = (B^T x_c)_k for x in reference order (nchains x N): block (k, c)
// walks column k's entries (padded indices, converted) with a fixed
// reduction order.
__global__ void __launch_bounds__(256) k_bt_ref(const double* __restrict__ x, long long N,
                                                const int32_t* __restrict__ bptr, const long long* __restrict__ bidx,
                                                const double* __restrict__ bval, DevGeom g, int m,
                                                double* __restrict__ z) {
  const int kc = blockIdx.x, c = blockIdx.y;
  const double* xc = x + (long long)c * N;
  double s = 0.0;
  for (int e = bptr[kc] + threadIdx.x; e < bptr[kc + 1]; e += blockDim.x) s += bval[e] * xc[padded_to_ref(bidx[e], g)];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) z[(long long)c * m + kc] = red[0] + red[1] + red[2] + red[3];
}

// y_c[row] = sum_e val_e w_c[col_e] over the rows of B (padded row indices,
// converted), for w (nchains x m); y (nchains x N in reference order) is
// zero elsewhere (cleared before the launch).
__global__ void __launch_bounds__(256) k_b_ref(const double* __restrict__ w, const long long* __restrict__ rows,
                                               const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                               const double* __restrict__ val, int nrows, DevGeom g, long long N,
                                               int m, double* __restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= nrows) return;
  double s = 0.0;
  for (int e = ptr[r]; e < ptr[r + 1]; ++e) s += val[e] * w[(long long)c * m + col[e]];
  y[(long long)c * N + padded_to_ref(rows[r], g)] = s;
}

// The contraction of k_contract on vectors in the reference vertex order
// (interior vertices only, x fastest; src/lattice/lattice3d.hh:132-141), as a
// caller's device tensors hold them: each (block, chain) walks a grid-stride
// range of vertices, applies S with Dirichlet boundaries (neighbours outside
// the interior are zero) and writes its two partial sums to
// part[(c * gridDim.x + block) * 2 + {0, 1}]. Every partial is overwritten.
__global__ void __launch_bounds__(256) k_contract_ref(const double* __restrict__ lam, const double* __restrict__ mu,
                                                      int dim, int nx1, int ny1, int nz1, long long N,
                                                      StencilT<double> S, double* __restrict__ part) {
  const int c = blockIdx.y;
  const double* l = lam + (long long)c * N;
  const double* u = mu + (long long)c * N;
  const long long plane = (long long)nx1 * ny1;
  const int zr = dim == 3 ? 1 : 0;
  double s0 = 0.0, s1 = 0.0;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(v % nx1);
    const long long r = v / nx1;
    const int j = (int)(r % ny1);
    const int k = (int)(r / ny1);
    double acc = 0.0;
    for (int dz = -zr; dz <= zr; ++dz) {
      if (k + dz < 0 || k + dz >= nz1) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        if (j + dy < 0 || j + dy >= ny1) continue;
        const double* row = u + v + dz * plane + dy * nx1;
        const int cb = (dim == 3 ? (dz + 1) * 9 : 0) + (dy + 1) * 3;
        if (i > 0) acc += S.c[cb] * row[-1];
        acc += S.c[cb + 1] * row[0];
        if (i + 1 < nx1) acc += S.c[cb + 2] * row[1];
      }
    }
    s0 += l[v] * u[v];
    s1 += l[v] * acc;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  __shared__ double red[2][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = s0;
    red[1][wid] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = part + ((long long)c * gridDim.x + blockIdx.x) * 2;
    o[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    o[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// out[2c + {0, 1}] = sum over the nblk partials of chain c (one wavefront per
// chain, lanes take partials lane, lane + 64, ... in order, then a fixed
// shuffle tree: the same result on every call).
__global__ void __launch_bounds__(64) k_contract_finish(const double* __restrict__ part, int nblk,
                                                        double* __restrict__ out) {
  const int c = blockIdx.x, lane = threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  for (int b = lane; b < nblk; b += 64) {
    s0 += part[((long long)c * nblk + b) * 2];
    s1 += part[((long long)c * nblk + b) * 2 + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  if (lane == 0) {
    out[2 * c] = s0;
    out[2 * c + 1] = s1;
  }
}

}  // namespace mgmc
