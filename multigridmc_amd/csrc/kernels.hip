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
      const T a0 = (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + 0];
      const T a1 = (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + 1];
      const T a2 = (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + 2];
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int q = 2 * l + IPAR;
        acc[l] += a0 * w[q] + a1 * w[q + 1] + a2 * w[q + 2];
      }
    }
  }
  const T om_d = (T)st.inv_diag_omega;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int q = 2 * l + IPAR;
    const int i = i0 + q;
    if (i < 1 || i > g.n0 - 1) continue;
    T b = fb[q];
    if (NOISE == NOISE_PHILOX) b += (T)st.noise_scale * (T)xi[l];
    if (NOISE == NOISE_EXTERNAL) b += (T)st.noise_scale * extg[q];
    xc[q + 1] += om_d * (b - acc[l]);
  }
}

// ---------------------------------------------------------------------------
// One colour phase of an SOR sweep (SORSmoother::apply_sparse,
// src/smoother/sor_smoother.cc:56-77, restricted to one colour), with the
// Gibbs right-hand side c = f + sqrt(D(2-omega)/omega) xi of
// SORSampler::apply (src/sampler/sor_sampler.cc:42-48) when NOISE != 0.
// STAR: colour = (i+j+k)&1; box: colour = (i&1) | (j&1)<<1 | (k&1)<<2.
// ---------------------------------------------------------------------------
template <typename T, int DIM, bool STAR, int NOISE>
__global__ void __launch_bounds__(256) k_sweep(T* __restrict__ x, const T* __restrict__ f,
                                               const T* __restrict__ ext, DevGeom g, Stencil st, int colour,
                                               int nrows, int K, NoiseKey nk) {
  const int ngx = g.px >> 3;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int gx = (int)(t % ngx);
  long long row = t / ngx;
  if (row >= (long long)nrows * K) return;
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
  load8<T>(f + base, fb);

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

// ---------------------------------------------------------------------------
// Low-rank (measurement) terms.
// z[c][k] = sum_e B_ek x_c[idx_e]   (B^T x, FP64 accumulation)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_lr_dot(const T* __restrict__ x, long long cs, const int32_t* __restrict__ bptr,
                                                const long long* __restrict__ bidx, const double* __restrict__ bval,
                                                int m, double* __restrict__ z) {
  int kcol = blockIdx.x, c = blockIdx.y;
  const T* xc = x + c * cs;
  double s = 0.0;
  for (int e = bptr[kcol] + threadIdx.x; e < bptr[kcol + 1]; e += blockDim.x) s += bval[e] * (double)xc[bidx[e]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[4];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) z[(long long)c * m + kcol] = red[0] + red[1] + red[2] + red[3];
}

// x_c[row_r] += sum_e val_e u_c[col_e] with u_c = M (a sqrt(Sigma) xi'_c + zc z_c).
// Smoother:   M = S^{-1}, a = 0, zc = -1  ->  x -= W S^{-1} B^T x = B_bar B^T x
//             (src/smoother/sor_smoother.cc:46-51).
// Sampler:    M = S^{-1}, a = 1, zc = -1  ->  the SMW step with the sampler's
//             low-rank noise B Sigma^{-1/2} xi' folded in
//             (src/sampler/sor_sampler.cc:50-56): x' = y + W S^{-1} (Sigma^{1/2} xi' - B^T y).
// Residual:   M = Sigma^{-1}, a = 0, zc = -1 -> subtract B Sigma^{-1} B^T x.
constexpr int MMAX = 128;
template <typename T, int NOISE>
__global__ void __launch_bounds__(256) k_lr_update(T* __restrict__ x, long long cs, const long long* __restrict__ rows,
                                                   const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                   const double* __restrict__ val, int nrows, const double* __restrict__ M,
                                                   const double* __restrict__ sqrt_sigma, const double* __restrict__ z,
                                                   double zc, int m, const double* __restrict__ ext_lr, NoiseKey nk) {
  __shared__ double a[MMAX];
  __shared__ double u[MMAX];
  int c = blockIdx.y;
  int tid = threadIdx.x;
  if (tid < m) {
    double v = zc * z[(long long)c * m + tid];
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
    double s = 0.0;
    for (int l = 0; l < m; ++l) s += M[tid * m + l] * a[l];
    u[tid] = s;
  }
  __syncthreads();
  int r = blockIdx.x * blockDim.x + tid;
  if (r >= nrows) return;
  double s = 0.0;
  for (int e = ptr[r]; e < ptr[r + 1]; ++e) s += val[e] * u[col[e]];
  T* p = x + c * cs + rows[r];
  *p = (T)((double)*p + s);
}

// ---------------------------------------------------------------------------
// r = f - A x on interior vertices (the sparse part of
// src/linear_operator/linear_operator.hh:78-88, multigridmc_sampler.cc:125-126)
// ---------------------------------------------------------------------------
template <typename T, int DIM, bool STAR>
__global__ void __launch_bounds__(256) k_residual(const T* __restrict__ x, const T* __restrict__ f, T* __restrict__ r,
                                                  DevGeom g, Stencil st, int K) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long per = (long long)(g.n0 - 1) * (g.n1 - 1) * (DIM == 3 ? g.n2 - 1 : 1);
  if (t >= per * K) return;
  int c = (int)(t / per);
  long long e = t % per;
  int i = (int)(e % (g.n0 - 1)) + 1;
  e /= (g.n0 - 1);
  int j = (int)(e % (g.n1 - 1)) + 1;
  int k = DIM == 3 ? (int)(e / (g.n1 - 1)) + 1 : 0;
  const long long pl = (long long)g.px * g.py;
  const long long p = c * g.cs + ((long long)k * g.py + j) * g.px + i;
  T acc;
  if (STAR) {
    // stencil index (dz+1)*9 + (dy+1)*3 + (dx+1): x-+ 12/14, y-+ 10/16, z-+ 4/22
    acc = (T)st.c[10] * x[p - g.px] + (T)st.c[12] * x[p - 1] + (T)st.diag * x[p] + (T)st.c[14] * x[p + 1] +
          (T)st.c[16] * x[p + g.px];
    if (DIM == 3) acc = (T)st.c[4] * x[p - pl] + acc + (T)st.c[22] * x[p + pl];
  } else {
    acc = (T)0;
#pragma unroll
    for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx)
          acc += (T)st.c[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] * x[p + dz * pl + dy * g.px + dx];
  }
  r[p] = (f ? f[p] : (T)0) - acc;
}

// fc[I] = sum_sigma p(sigma) r[2I + sigma], p = prod {1/2, 1, 1/2}
// (IntergridOperator::restrict, src/intergrid/intergrid_operator.hh:74-88)
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_restrict(const T* __restrict__ r, T* __restrict__ fc, DevGeom gf, DevGeom gc,
                                                  int K) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long per = (long long)(gc.n0 - 1) * (gc.n1 - 1) * (DIM == 3 ? gc.n2 - 1 : 1);
  if (t >= per * K) return;
  int c = (int)(t / per);
  long long e = t % per;
  int I = (int)(e % (gc.n0 - 1)) + 1;
  e /= (gc.n0 - 1);
  int J = (int)(e % (gc.n1 - 1)) + 1;
  int Kz = DIM == 3 ? (int)(e / (gc.n1 - 1)) + 1 : 0;
  const long long pl = (long long)gf.px * gf.py;
  const long long p = c * gf.cs + ((long long)(2 * Kz) * gf.py + 2 * J) * gf.px + 2 * I;
  const T w[3] = {(T)0.5, (T)1.0, (T)0.5};
  T acc = (T)0;
#pragma unroll
  for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const T* q = r + p + dz * pl + dy * gf.px;
      T wz = DIM == 3 ? w[dz + 1] : (T)1;
      acc += wz * w[dy + 1] * ((T)0.5 * q[-1] + q[0] + (T)0.5 * q[1]);
    }
  fc[c * gc.cs + ((long long)Kz * gc.py + J) * gc.px + I] = acc;
}

// x += alpha P xc (IntergridOperator::prolongate_add in gather form,
// src/intergrid/intergrid_operator.hh:106-120)
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_prolongate(T* __restrict__ x, const T* __restrict__ xc, DevGeom gf, DevGeom gc,
                                                    T alpha, int K) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long per = (long long)(gf.n0 - 1) * (gf.n1 - 1) * (DIM == 3 ? gf.n2 - 1 : 1);
  if (t >= per * K) return;
  int c = (int)(t / per);
  long long e = t % per;
  int i = (int)(e % (gf.n0 - 1)) + 1;
  e /= (gf.n0 - 1);
  int j = (int)(e % (gf.n1 - 1)) + 1;
  int k = DIM == 3 ? (int)(e / (gf.n1 - 1)) + 1 : 0;
  int i0 = i >> 1, i1 = (i + 1) >> 1, j0 = j >> 1, j1 = (j + 1) >> 1, k0 = k >> 1, k1 = (k + 1) >> 1;
  T wi = (i & 1) ? (T)0.5 : (T)1, wj = (j & 1) ? (T)0.5 : (T)1, wk = (DIM == 3 && (k & 1)) ? (T)0.5 : (T)1;
  const T* b = xc + c * gc.cs;
  auto at = [&](int a, int bb, int cc) { return b[((long long)cc * gc.py + bb) * gc.px + a]; };
  T s;
  if (DIM == 3) {
    T s0 = at(i0, j0, k0) + ((i & 1) ? at(i1, j0, k0) : (T)0);
    T s1 = (j & 1) ? at(i0, j1, k0) + ((i & 1) ? at(i1, j1, k0) : (T)0) : (T)0;
    T sa = s0 + s1;
    T sb = (T)0;
    if (k & 1) {
      T s2 = at(i0, j0, k1) + ((i & 1) ? at(i1, j0, k1) : (T)0);
      T s3 = (j & 1) ? at(i0, j1, k1) + ((i & 1) ? at(i1, j1, k1) : (T)0) : (T)0;
      sb = s2 + s3;
    }
    s = wi * wj * wk * (sa + sb);
  } else {
    T s0 = at(i0, j0, 0) + ((i & 1) ? at(i1, j0, 0) : (T)0);
    T s1 = (j & 1) ? at(i0, j1, 0) + ((i & 1) ? at(i1, j1, 0) : (T)0) : (T)0;
    s = wi * wj * (s0 + s1);
  }
  x[c * gf.cs + ((long long)k * gf.py + j) * gf.px + i] += alpha * s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_copy(T* __restrict__ dst, const T* __restrict__ src, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ p, long long n, T v) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

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
// (src/solver/cholesky_solver.cc:30-40); FP64 inside. One block per chain.
template <typename T, int NOISE>
__global__ void __launch_bounds__(256) k_coarse(T* __restrict__ x, const T* __restrict__ f, long long cs,
                                                const long long* __restrict__ pidx, const double* __restrict__ LinvT,
                                                const double* __restrict__ Qinv, int Nc, const double* __restrict__ ext,
                                                NoiseKey nk) {
  MGMC_DYNAMIC_SHARED(double, sh);
  double* fs = sh;
  double* xs = sh + Nc;
  int c = blockIdx.x;
  for (int j = threadIdx.x; j < Nc; j += blockDim.x) {
    fs[j] = (double)f[c * cs + pidx[j]];
    if (NOISE == NOISE_PHILOX) {
      float xi[4];
      normals4(COARSE_SLOT + (uint32_t)(j >> 2), (uint32_t)c, nk.tag, *nk.sample, nk.k0, nk.k1, xi);
      xs[j] = (double)xi[j & 3];
    } else if (NOISE == NOISE_EXTERNAL) {
      xs[j] = ext[(long long)c * Nc + j];
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < Nc; r += blockDim.x) {
    double s = 0.0;
    const double* q = Qinv + (long long)r * Nc;
    for (int j = 0; j < Nc; ++j) s += q[j] * fs[j];
    if (NOISE != NOISE_NONE) {
      const double* l = LinvT + (long long)r * Nc;
      for (int j = r; j < Nc; ++j) s += l[j] * xs[j];
    }
    x[c * cs + pidx[r]] = (T)s;
  }
}

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
template <typename T, int NOISE>
__global__ void __launch_bounds__(1024) k_sweep_rb3(const T* __restrict__ x, T* __restrict__ xout,
                                                    const T* __restrict__ f, const T* __restrict__ ext, DevGeom g,
                                                    Stencil st, int first, int TY, int SZ, int ntile, int nseg,
                                                    NoiseKey nk) {
  MGMC_DYNAMIC_SHARED(char, lds_raw);
  const size_t asz = (size_t)(TY + 4) * g.px, bsz = (size_t)(TY + 2) * g.px;
  T* A0 = reinterpret_cast<T*>(lds_raw);
  T* B0 = A0 + 2 * asz;
  const int ngx = g.px >> 3;
  const int tid = threadIdx.x;
  const int r = tid / ngx, gx = tid % ngx;
  int b = blockIdx.x;
  const int tile = b % ntile;
  b /= ntile;
  const int seg = b % nseg;
  const int chain = b / nseg;
  const int j0 = 1 + tile * TY;
  const int j = j0 - 2 + r;
  const bool active = r < TY + 4;
  const bool has_mem = active && j >= 0 && j <= g.n1;
  const bool ring = active && j >= j0 - 1 && j <= j0 + TY && j >= 1 && j <= g.n1 - 1;
  const bool owned = active && j >= j0 && j <= j0 + TY - 1 && j <= g.n1 - 1;
  const int ks = 1 + seg * SZ;
  const int ke = ks + SZ < g.n2 ? ks + SZ : g.n2;  // owned interior planes [ks, ke)
  const long long pl = (long long)g.px * g.py;
  const long long cbase = (long long)chain * g.cs + (long long)j * g.px + 8 * gx;
  const uint32_t sample = *nk.sample;
  const int second = 1 - first;

  T xmm[8], xm[8], xc[8], xp[8], fm[8], fc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) xmm[q] = xm[q] = xc[q] = xp[q] = fm[q] = fc[q] = (T)0;
  if (has_mem && ks - 2 >= 0) load8<T>(x + cbase + (ks - 2) * pl, xm);
  if (has_mem) load8<T>(x + cbase + (ks - 1) * pl, xc);

  const T cx = (T)st.c[12], cy = (T)st.c[10], cz = (T)st.c[4], D = (T)st.diag;
  const T om_d = (T)st.inv_diag_omega, ns = (T)st.noise_scale;
  const size_t aoff = (size_t)r * g.px + 8 * gx;        // A rows j0-2 .. j0+TY+1
  const size_t boff = (size_t)(r - 1) * g.px + 8 * gx;  // B rows j0-1 .. j0+TY (r = 1 .. TY+2)

  for (int k = ks - 1; k <= ke; ++k) {
    const int kb = k - 1;
    const bool doA = ring && k >= 1 && k <= g.n2 - 1;
    const bool doB = owned && kb >= ks && kb <= ke - 1;
    // loads first: next plane of x, f on plane k
    if (has_mem && k + 1 <= g.n2) load8<T>(x + cbase + (long long)(k + 1) * pl, xp);
    else {
#pragma unroll
      for (int q = 0; q < 8; ++q) xp[q] = (T)0;
    }
    if (doA) load8<T>(f + cbase + (long long)k * pl, fc);
    // noise of both colours while the loads are in flight
    const int iparA = (first + j + k) & 1, iparB = (second + j + kb) & 1;
    float xiA[4] = {0.f, 0.f, 0.f, 0.f}, xiB[4] = {0.f, 0.f, 0.f, 0.f};
    if (NOISE == NOISE_PHILOX) {
      if (doA) {
        uint32_t c0 = (uint32_t)((((long long)k * g.py + j) * ngx + gx) << 1) | (uint32_t)iparA;
        normals4(c0, (uint32_t)chain, nk.tag, sample, nk.k0, nk.k1, xiA);
      }
      if (doB) {
        uint32_t c0 = (uint32_t)((((long long)kb * g.py + j) * ngx + gx) << 1) | (uint32_t)iparB;
        normals4(c0, (uint32_t)chain, nk.tag, sample, nk.k0, nk.k1, xiB);
      }
    }
    T* A = A0 + (k & 1) * asz;         // plane k before the first colour
    const T* Bm = B0 + (kb & 1) * bsz;  // plane k-1 after the first colour (written last step)
    T* Bk = B0 + (k & 1) * bsz;
    if (active) store8<T>(A + aoff, xc);
    __syncthreads();
    // first colour on plane k (owned rows and halo ring)
    if (doA) {
      const T* Arow = A + aoff;
      T ym[8], yp[8];
      load8<T>(Arow - g.px, ym);
      load8<T>(Arow + g.px, yp);
      const T left = gx > 0 ? Arow[-1] : (T)0;
      const T right = gx < ngx - 1 ? Arow[8] : (T)0;
      const T* eg = NOISE == NOISE_EXTERNAL ? ext + cbase + (long long)k * pl : nullptr;
      if (iparA == 0)
        star_update4<0, T, NOISE>(xc, left, right, ym, yp, xm, xp, fc, xiA, eg, 8 * gx, g.n0, cx, cy, cz, D, om_d, ns);
      else
        star_update4<1, T, NOISE>(xc, left, right, ym, yp, xm, xp, fc, xiA, eg, 8 * gx, g.n0, cx, cy, cz, D, om_d, ns);
    }
    // second colour on plane k-1 (owned rows, owned planes)
    if (doB) {
      const T* Brow = Bm + boff;
      T ym[8], yp[8];
      load8<T>(Brow - g.px, ym);
      load8<T>(Brow + g.px, yp);
      const T left = gx > 0 ? Brow[-1] : (T)0;
      const T right = gx < ngx - 1 ? Brow[8] : (T)0;
      const T* eg = NOISE == NOISE_EXTERNAL ? ext + cbase + (long long)kb * pl : nullptr;
      if (iparB == 0)
        star_update4<0, T, NOISE>(xm, left, right, ym, yp, xmm, xc, fm, xiB, eg, 8 * gx, g.n0, cx, cy, cz, D, om_d, ns);
      else
        star_update4<1, T, NOISE>(xm, left, right, ym, yp, xmm, xc, fm, xiB, eg, 8 * gx, g.n0, cx, cy, cz, D, om_d, ns);
      store8<T>(xout + cbase + (long long)kb * pl, xm);
    }
    // plane k after the first colour, for the second colour at the next step
    if (active && r >= 1 && r <= TY + 2) store8<T>(Bk + boff, xc);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      xmm[q] = xm[q];
      xm[q] = xc[q];
      xc[q] = xp[q];
      fm[q] = fc[q];
    }
  }
}

// obs[row][c][p] = x_c[probe_p], row taken from the device run counter
template <typename T>
__global__ void k_probe(const T* __restrict__ x, long long cs, const long long* __restrict__ pidx, int P, int K,
                        double* __restrict__ obs, const uint32_t* __restrict__ ctr) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * K) return;
  int c = t / P, p = t % P;
  uint32_t row = ctr[1];
  obs[((long long)row * K + c) * P + p] = (double)x[c * cs + pidx[p]];
}

// advance the sample counter and the observable row after a cycle
__global__ void k_advance(uint32_t* ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctr[0] += 1u;
    ctr[1] += 1u;
  }
}

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
__global__ void k_sumsq(const double* __restrict__ x, long long n, double* __restrict__ out) {
  double s = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    s += x[t] * x[t];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double red[4];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace mgmc
