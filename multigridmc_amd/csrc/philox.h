// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller
// normals. Written from the published round function; usable on host and
// device so that the host-side checks see exactly the integer stream the
// kernels use.
//
// Noise addressing. Every Gaussian the sampler needs is a pure function of
// (seed, sample counter, tag, chain, slot):
//   * tag numbers the noise-consuming launches inside one MGMC cycle,
//   * a lattice vertex (i, j, k) of a padded level with pitch px (multiple of
//     8) and row id r = k * py + j uses slot
//       ctr0 = ((r * (px / 8) + (i >> 3)) << 1) | (i & 1),  lane = (i & 7) >> 1,
//     so the four vertices of equal x-parity in an aligned group of eight
//     share one Philox call (one call per four normals),
//   * low-rank noise xi'_k uses ctr0 = LOWRANK_SLOT + (k >> 2), lane k & 3,
//   * coarsest-level noise xi_j uses ctr0 = COARSE_SLOT + (j >> 2), lane j & 3.
// The counter is (ctr0, chain, tag, sample), the key is the 64-bit seed.
#ifndef MGMC_PHILOX_H
#define MGMC_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MGMC_HD __host__ __device__ __forceinline__
#else
#define MGMC_HD inline
#endif

namespace mgmc {

constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;  // golden ratio
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;  // sqrt(3) - 1

constexpr uint32_t LOWRANK_SLOT = 0xC0000000u;
constexpr uint32_t COARSE_SLOT = 0xD0000000u;

struct u32x4 {
  uint32_t v[4];
};

MGMC_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// a ^ b ^ c (one v_bitop3_b32 on the device: the compiler does not fuse
// the two XORs of a Philox round by itself)
MGMC_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

MGMC_HD u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(PHILOX_M0, c0, hi0, lo0);
    mulhilo(PHILOX_M1, c2, hi1, lo1);
    uint32_t n0 = xor3(hi1, c1, k0);
    uint32_t n2 = xor3(hi0, c3, k1);
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  u32x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// Uniform in (0, 1): the top 24 bits, centred in their cell.
MGMC_HD float u24(uint32_t b) { return (float)(b >> 8) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f); }

}  // namespace mgmc

#endif  // MGMC_PHILOX_H
