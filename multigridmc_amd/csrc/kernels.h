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

#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

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

// The workgroup a kernel body works on: its block coordinates and the grid
// of its own launch. A plain launch passes the hardware's (Blk::hw()); a
// merged launch of several instances (k_multi) passes each instance's own.
struct Blk {
  uint32_t x, y, gx, gy;
  __device__ static Blk hw() { return Blk{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}; }
};

// ---------------------------------------------------------------------------
// Merged launches. A kernel whose body is a __device__ function of a Blk and
// its parameters is "batchable": the launches that several independent
// instances (lattices of different sizes, each with its own fields and
// hierarchy) make of it at the same point of the cycle become one launch of
// k_multi, which reads the instance's arguments from a table. Item i owns the
// merged blocks [start_i, start_i + gx_i gy_i); starts are multiples of 8, so
// that a block keeps the XCD (block index mod 8) it has in a plain launch.
// With fewer workgroups than blocks (`total`), the workgroups loop over the
// blocks (persistent form); bodies then must reach every barrier on all
// threads, which holds for every batchable body of this library.
// ---------------------------------------------------------------------------
template <typename Tup>
struct MultiItem {
  uint32_t start, gx, gy, pad;
  Tup args;
};
template <auto Body, typename Tup, size_t... I>
__device__ __forceinline__ void multi_call(const Blk& bk, const Tup& t, std::index_sequence<I...>) {
  Body(bk, std::get<I>(t)...);
}
template <auto Body, int MAXT, typename Tup>
__global__ void __launch_bounds__(MAXT) k_multi(const MultiItem<Tup>* __restrict__ items, int n, uint32_t total) {
  for (uint32_t b = blockIdx.x; b < total; b += gridDim.x) {
    if (b != blockIdx.x) __syncthreads();  // the previous block's LDS is consumed
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].start <= b) lo = mid;
      else hi = mid - 1;
    }
    const MultiItem<Tup>* it = items + lo;
    const uint32_t local = b - it->start;
    if (local >= it->gx * it->gy) continue;  // alignment padding between items
    const Blk bk{local % it->gx, local / it->gx, it->gx, it->gy};
    multi_call<Body>(bk, it->args, std::make_index_sequence<std::tuple_size<Tup>::value>{});
  }
}

// A launch recorded for merging: the k_multi kernel of its body, its shape
// and one MultiItem (start 0) holding its arguments.
using MultiLauncher = void (*)(dim3 grid, dim3 block, size_t shmem, hipStream_t stream, const void* items, int n,
                               uint32_t total);
struct LaunchRec {
  const void* multi = nullptr;  // k_multi instantiation (identifies the body)
  MultiLauncher launcher = nullptr;
  bool tolerant = false;  // the body computes the same with more threads per block than asked for
  dim3 grid, block;
  size_t shmem = 0;
  std::vector<char> item;
  size_t item_align = 8;
  int dir = 0, level = 0, stage = 0, seq = 0, instance = 0;
};
// Launches made while a recorder is installed (on this thread) are recorded
// instead of issued; dir / level / stage say where in the cycle they are
// (set by the engine), seq numbers launches within one (dir, level, stage).
struct Recorder {
  std::vector<LaunchRec> recs;
  int dir = 0, level = 0, stage = 0;
  int instance = 0;
};
extern thread_local Recorder* g_rec;
inline void rec_mark(int dir, int level, int stage) {
  if (g_rec) {
    g_rec->dir = dir;
    g_rec->level = level;
    g_rec->stage = stage;
  }
}
[[noreturn]] void unbatchable_launch();

// Kernel launch through hipLaunchKernel with each argument converted to the
// kernel's parameter type (one entry point for every launch of the engine).
template <typename Tuple, size_t... I>
inline void launch_args(Tuple& t, void** argv, std::index_sequence<I...>) {
  ((argv[I] = (void*)&std::get<I>(t)), ...);
}
template <typename... P, typename... A>
inline void launch(dim3 grid, dim3 block, size_t shmem, hipStream_t stream, void (*k)(P...), A&&... args) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  if (g_rec) unbatchable_launch();
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

// A batchable kernel: its body, the workgroup size bound of its kernels,
// whether the body tolerates larger workgroups, and its plain kernel (which
// calls the body with Blk::hw()). Engines launch through these tags.
template <auto Body, int MAXT, bool TOL, auto Kern>
struct Batchable {};

template <auto Body, int MAXT, typename Tup>
void multi_launch(dim3 grid, dim3 block, size_t shmem, hipStream_t stream, const void* items, int n, uint32_t total) {
  auto saved = g_rec;
  g_rec = nullptr;
  launch(grid, block, shmem, stream, k_multi<Body, MAXT, Tup>, static_cast<const MultiItem<Tup>*>(items), n, total);
  g_rec = saved;
}
template <auto Body, int MAXT, bool TOL, typename... P, typename... A>
void record_launch(void (*)(P...), dim3 grid, dim3 block, size_t shmem, A&&... args) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  using Tup = std::tuple<std::decay_t<P>...>;
  MultiItem<Tup> it{0u, grid.x, grid.y, 0u, Tup(static_cast<P>(args)...)};
  LaunchRec r;
  r.multi = reinterpret_cast<const void*>(k_multi<Body, MAXT, Tup>);
  r.launcher = &multi_launch<Body, MAXT, Tup>;
  r.tolerant = TOL;
  r.grid = grid;
  r.block = block;
  r.shmem = shmem;
  r.item.resize(sizeof(it));
  std::memcpy(r.item.data(), static_cast<const void*>(&it), sizeof(it));
  r.item_align = alignof(MultiItem<Tup>);
  Recorder& R = *g_rec;
  r.dir = R.dir;
  r.level = R.level;
  r.stage = R.stage;
  r.instance = R.instance;
  r.seq = 0;
  for (const LaunchRec& q : R.recs)
    if (q.dir == r.dir && q.level == r.level && q.stage == r.stage) ++r.seq;
  R.recs.push_back(std::move(r));
}
template <auto Body, int MAXT, bool TOL, auto Kern, typename... A>
inline void launch(dim3 grid, dim3 block, size_t shmem, hipStream_t stream, Batchable<Body, MAXT, TOL, Kern>,
                   A&&... args) {
  if (g_rec) {
    if (grid.z != 1) unbatchable_launch();
    record_launch<Body, MAXT, TOL>(Kern, grid, block, shmem, std::forward<A>(args)...);
    return;
  }
  launch(grid, block, shmem, stream, Kern, std::forward<A>(args)...);
}

// Four N(0,1) normals of one Philox call (Box-Muller on two uniform pairs).
// v_log_f32 is log2 and v_sin_f32 / v_cos_f32 take revolutions, so
// sin(2 pi u) needs no range reduction for u in [0, 1).
// Makes v look changed to the compiler (no instruction is emitted): inside
// a loop, values derived from it (the Philox key schedule) are then
// recomputed with scalar adds instead of being held in scalar registers,
// which the level-0 sweep otherwise spills into vector-register lanes
// (128^3 x 64 FP32 chains: sweep -0.55 %, same-box A/B).
__device__ __forceinline__ void opaque_sgpr(uint32_t& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+s"(v));
#else
  (void)v;
#endif
}

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
