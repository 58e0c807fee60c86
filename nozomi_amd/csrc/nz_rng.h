// Counter-based random numbers for the walker: Philox4x32-10 (Salmon et al.,
// SC'11) keyed on the 64-bit seed, counter (draw, step, spin, stream).
// Stateless, so any spin's stream can be regenerated from its coordinates:
// results are the same whatever the launch shape, batch or step chunking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nz {

enum : uint32_t { kStreamWalk = 0u, kStreamSeed = 1u };

struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
    const uint32_t lo1 = 0xCD9E8D57u * c.z;
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// ((x >> 9) + 1/2) * 2^-23: 23-bit uniform on (0,1), exact in binary32.
__device__ __forceinline__ float unif23(uint32_t x) {
  return __uint2float_rn((x >> 9) << 1 | 1u) * 5.9604644775390625e-8f;  // * 2^-24
}

// Three standard normals (Box-Muller on (u0,u1) and (u2,u3)).
__device__ __forceinline__ void normals3(uint64_t seed, uint32_t stream, uint32_t spin,
                                         uint32_t step, uint32_t draw, float& n0, float& n1,
                                         float& n2) {
  const u32x4 r = philox4x32_10(u32x4{draw, step, spin, stream}, (uint32_t)seed,
                                (uint32_t)(seed >> 32));
  const float r0 = sqrtf(-2.0f * logf(unif23(r.x)));
  const float r1 = sqrtf(-2.0f * logf(unif23(r.z)));
  float s0, c0;
  sincospif(2.0f * unif23(r.y), &s0, &c0);
  n0 = r0 * c0;
  n1 = r0 * s0;
  n2 = r1 * cospif(2.0f * unif23(r.w));
}

}  // namespace nz
