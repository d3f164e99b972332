// CPU emulation of the HIP runtime subset libmgmc_amd uses, for the
// sanitizer tests: the library compiled by g++ with -DMGMC_EMULATION runs its
// kernels on the host, one fiber per GPU thread of a block (blocks one after
// another), so AddressSanitizer sees every device-memory and LDS access.
//
// - device memory is malloc'd (exact sizes, ASan red zones around each
//   allocation); dynamic LDS is a malloc'd buffer of the launch's size per
//   block; static __shared__ arrays are function-local statics (one block
//   runs at a time, so they are block-shared);
// - the threads of a block are fibers on the calling thread; __syncthreads
//   waits for every live thread of the block and __shfl_down for the live
//   lanes of the wave; a thread that returns leaves both;
// - streams and events are synchronous, stream capture is not supported (the
//   tests run cycles eagerly);
// - the device fast-math builtins map to libm: v_log_f32 is log2, v_sin_f32
//   and v_cos_f32 take revolutions.
// Test infrastructure only: nothing in the product includes it.
#ifndef MGMC_HIP_EMU_H
#define MGMC_HIP_EMU_H

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <tuple>
#include <ucontext.h>
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#endif
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3 {
  unsigned x, y, z;
};
inline uint3 blockIdx;
inline dim3 blockDim, gridDim;  // threadIdx: the running fiber's (below)

struct float4 {
  float x, y, z, w;
};
struct double2 {
  double x, y;
};
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
inline double2 make_double2(double a, double b) { return double2{a, b}; }

inline float __builtin_amdgcn_logf(float v) { return std::log2(v); }
inline float __builtin_amdgcn_sinf(float v) { return std::sin(6.283185307179586f * v); }
inline float __builtin_amdgcn_cosf(float v) { return std::cos(6.283185307179586f * v); }
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }  // the values it is used on are uniform

// ---- runtime API ----
typedef enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotSupported = 801 } hipError_t;
typedef enum {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
} hipMemcpyKind;
struct ihipStream_t {
  int id;
};
struct ihipEvent_t {
  std::chrono::steady_clock::time_point t;
};
struct ihipGraph_t {
  int id;
};
struct ihipGraphExec_t {
  int id;
};
typedef ihipStream_t* hipStream_t;
typedef ihipEvent_t* hipEvent_t;
typedef ihipGraph_t* hipGraph_t;
typedef ihipGraphExec_t* hipGraphExec_t;
enum { hipStreamNonBlocking = 1 };
typedef enum { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 } hipStreamCaptureMode;

inline const char* hipGetErrorString(hipError_t e) {
  return e == hipSuccess ? "no error" : e == hipErrorNotSupported ? "not supported in emulation" : "invalid value";
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
template <typename T>
inline hipError_t hipMalloc(T** p, size_t n) {
  *p = static_cast<T*>(std::malloc(n ? n : 1));
  return *p ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t = nullptr) {
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) {
  if (n) std::memset(d, v, n);
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new ihipStream_t{1};
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t{};
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
enum { hipEventDisableTiming = 2, hipEventDisableSystemFence = 0x20000000 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, char*, size_t) {
  return hipErrorNotSupported;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

// ---- kernel execution ----
// The threads of a block are fibers (ucontext) on the calling thread, run
// round-robin by a scheduler; __syncthreads and the wave barrier of
// __shfl_down yield to it, and it releases a barrier once every live fiber of
// the block (or wave) waits there. Switches are announced to AddressSanitizer.
namespace emu {

constexpr int kWave = 64;
constexpr size_t kStack = 64 * 1024;  // kernel frames are small; an overflow lands in the malloc red zone

enum State { RUNNABLE, WAIT_BLOCK, WAIT_WAVE, DONE };
struct Fiber {
  ucontext_t ctx;
  char* stack = nullptr;
  uint3 tidx;
  int tid = 0;
  State state = DONE;
};

struct Block {
  std::vector<Fiber> fibers;
  std::vector<uint64_t> slots;  // one 8-byte slot per thread for __shfl_down
  char* lds = nullptr;
  int nt = 0;
  std::function<void()> body;
  ucontext_t main_ctx;
  const void* main_bottom = nullptr;
  size_t main_size = 0;
};
inline Block g_block;
inline Fiber* g_cur = nullptr;

inline void switch_to(ucontext_t* from, ucontext_t* to, const void* bottom, size_t size, bool dying) {
#if defined(__SANITIZE_ADDRESS__)
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(dying ? nullptr : &fake, bottom, size);
  swapcontext(from, to);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
  (void)bottom;
  (void)size;
  (void)dying;
  swapcontext(from, to);
#endif
}

// from a fiber back to the scheduler
inline void yield(State s) {
  Fiber* f = g_cur;
  f->state = s;
  switch_to(&f->ctx, &g_block.main_ctx, g_block.main_bottom, g_block.main_size, false);
}

inline void fiber_entry(int i) {
#if defined(__SANITIZE_ADDRESS__)
  __sanitizer_finish_switch_fiber(nullptr, &g_block.main_bottom, &g_block.main_size);
#endif
  Fiber* f = &g_block.fibers[(size_t)i];
  g_cur = f;
  g_block.body();
  f->state = DONE;
  switch_to(&f->ctx, &g_block.main_ctx, g_block.main_bottom, g_block.main_size, true);
}

inline std::vector<char*>& stacks() {
  static std::vector<char*> s;
  return s;
}

inline void run_block(int nt) {
  Block& b = g_block;
  b.nt = nt;
  b.fibers.assign((size_t)nt, Fiber{});
  b.slots.assign((size_t)nt, 0);
  auto& st = stacks();
  while ((int)st.size() < nt) st.push_back(static_cast<char*>(std::malloc(kStack)));
  for (int i = 0; i < nt; ++i) {
    Fiber& f = b.fibers[(size_t)i];
    f.tid = i;
    f.tidx = uint3{i % blockDim.x, (i / blockDim.x) % blockDim.y, i / (blockDim.x * blockDim.y)};
    f.stack = st[(size_t)i];
    f.state = RUNNABLE;
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, reinterpret_cast<void (*)()>(fiber_entry), 1, i);
  }
  const int nw = (nt + kWave - 1) / kWave;
  for (;;) {
    bool ran = false;
    for (int i = 0; i < nt; ++i) {
      Fiber& f = b.fibers[(size_t)i];
      if (f.state != RUNNABLE) continue;
      ran = true;
      g_cur = &f;
      switch_to(&b.main_ctx, &f.ctx, f.stack, kStack, false);
    }
    int live = 0, at_block = 0;
    for (auto& f : b.fibers) {
      live += f.state != DONE;
      at_block += f.state == WAIT_BLOCK;
    }
    if (live == 0) break;
    bool released = false;
    if (at_block == live) {
      for (auto& f : b.fibers)
        if (f.state == WAIT_BLOCK) f.state = RUNNABLE;
      released = true;
    }
    for (int w = 0; w < nw; ++w) {
      int wl = 0, ww = 0;
      for (int i = w * kWave; i < std::min(nt, (w + 1) * kWave); ++i) {
        wl += b.fibers[(size_t)i].state != DONE;
        ww += b.fibers[(size_t)i].state == WAIT_WAVE;
      }
      if (wl > 0 && ww == wl) {
        for (int i = w * kWave; i < std::min(nt, (w + 1) * kWave); ++i) b.fibers[(size_t)i].state = RUNNABLE;
        released = true;
      }
    }
    if (!ran && !released) {
      std::fprintf(stderr, "hip_emu: deadlock (threads wait at different barriers)\n");
      std::abort();
    }
  }
}

inline char* dyn_shared() { return g_block.lds; }

}  // namespace emu

#define threadIdx (emu::g_cur->tidx)
#define MGMC_DYNAMIC_SHARED(type, name) type* name = reinterpret_cast<type*>(emu::dyn_shared())

inline void __syncthreads() { emu::yield(emu::WAIT_BLOCK); }

template <typename T>
inline T __shfl_down(T v, unsigned delta, int width = emu::kWave) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  emu::Block& b = emu::g_block;
  const int tid = emu::g_cur->tid;
  const int lane = tid % emu::kWave;
  std::memcpy(&b.slots[(size_t)tid], &v, sizeof(T));
  emu::yield(emu::WAIT_WAVE);
  T r = v;
  if ((lane % width) + (int)delta < width && tid + (int)delta < b.nt) std::memcpy(&r, &b.slots[(size_t)tid + delta], sizeof(T));
  emu::yield(emu::WAIT_WAVE);
  return r;
}

template <typename... P, typename... A>
void emu_launch(dim3 grid, dim3 block, size_t shmem, void (*k)(P...), A&&... args) {
  const int nt = (int)(block.x * block.y * block.z);
  if (nt <= 0 || nt > 1024) std::abort();
  // the arguments as the kernel's parameter types, copied once per launch
  std::tuple<std::decay_t<P>...> tup(static_cast<P>(args)...);
  gridDim = grid;
  blockDim = block;
  for (unsigned bz = 0; bz < grid.z; ++bz)
    for (unsigned by = 0; by < grid.y; ++by)
      for (unsigned bx = 0; bx < grid.x; ++bx) {
        blockIdx = uint3{bx, by, bz};
        emu::g_block.lds = shmem ? static_cast<char*>(std::malloc(shmem)) : nullptr;
        emu::g_block.body = [&]() { std::apply([&](auto... a) { k(a...); }, tup); };
        emu::run_block(nt);
        std::free(emu::g_block.lds);
        emu::g_block.lds = nullptr;
      }
}

#endif
