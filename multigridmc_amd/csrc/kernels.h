// gfx950 kernels of the MGMC cycle. All fields use the padded per-level
// layout described in setup.h; `cs` is the per-chain stride and the chain
// index is the slowest coordinate of every launch.
#ifndef MGMC_KERNELS_H
#define MGMC_KERNELS_H

#ifdef MGMC_EMULATION
#include "hip_emu.h"  // host emulation of the sanitizer tests (tests/emu)
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include <tuple>
#include <type_traits>
#include <utility>

#include "philox.h"

namespace mgmc {

[[noreturn]] void launch_failed(hipError_t e);

struct DevGeom {
  int dim, n0, n1, n2;
  int px, py, pz;
  long long cs;   // chain stride
  long long fcs;  // chain stride of the right-hand side f (0: one f shared by all chains)
};

// Constant stencil (27 entries, (dz+1)*9 + (dy+1)*3 + (dx+1)) and the
// per-level SOR constants: omega/D and sqrt(D (2 - omega) / omega)
// (src/sampler/sor_sampler.cc:23-27).
struct Stencil {
  double c[27];
  double diag;
  double omega;
  double inv_diag_omega;  // omega / D
  double noise_scale;     // sqrt(D (2 - omega) / omega)
};

// The stencil and SOR constants in the field precision, for kernels whose
// coefficients then stay in scalar registers (no per-use conversion).
template <typename T>
struct StencilT {
  T c[27];
  T inv_diag_omega;
  T noise_scale;
};

// Philox key and the noise tag of a launch; the sample counter is read from
// device memory so that captured graphs can be replayed.
struct NoiseKey {
  uint32_t k0, k1;
  uint32_t tag;
  const uint32_t* sample;
};

#ifndef MGMC_DYNAMIC_SHARED
#define MGMC_DYNAMIC_SHARED(type, name) extern __shared__ type name[]
#endif

enum NoiseMode { NOISE_NONE = 0, NOISE_PHILOX = 1, NOISE_EXTERNAL = 2 };

// Kernel launch through hipLaunchKernel with each argument converted to the
// kernel's parameter type (one entry point for every launch of the engine).
template <typename Tuple, size_t... I>
inline void launch_args(Tuple& t, void** argv, std::index_sequence<I...>) {
  ((argv[I] = (void*)&std::get<I>(t)), ...);
}
template <typename... P, typename... A>
inline void launch(dim3 grid, dim3 block, size_t shmem, hipStream_t stream, void (*k)(P...), A&&... args) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
#ifdef MGMC_EMULATION
  (void)stream;
  emu_launch(grid, block, shmem, k, std::forward<A>(args)...);
#else
  std::tuple<std::decay_t<P>...> t(static_cast<P>(args)...);
  void* argv[sizeof...(P) + 1];
  launch_args(t, argv, std::index_sequence_for<P...>{});
  hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(k), grid, block, argv, shmem, stream);
  if (e != hipSuccess) launch_failed(e);
#endif
}

// Four N(0,1) normals of one Philox call (Box-Muller on two uniform pairs).
// v_log_f32 is log2 and v_sin_f32 / v_cos_f32 take revolutions, so
// sin(2 pi u) needs no range reduction for u in [0, 1).
__device__ __forceinline__ void normals4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                         uint32_t k1, float out[4]) {
  u32x4 r = philox4x32_10(c0, c1, c2, c3, k0, k1);
  const float m2ln2 = -2.0f * 0.69314718055994531f;
  float ra = __builtin_sqrtf(m2ln2 * __builtin_amdgcn_logf(u24(r.v[0])));
  float rb = __builtin_sqrtf(m2ln2 * __builtin_amdgcn_logf(u24(r.v[2])));
  float ua = u24(r.v[1]), ub = u24(r.v[3]);
  out[0] = ra * __builtin_amdgcn_cosf(ua);
  out[1] = ra * __builtin_amdgcn_sinf(ua);
  out[2] = rb * __builtin_amdgcn_cosf(ub);
  out[3] = rb * __builtin_amdgcn_sinf(ub);
}

}  // namespace mgmc

#endif
