// Counter-based random numbers for the walker: Philox4x32-10 (Salmon et al.,
// SC'11) keyed on the 64-bit seed, counter (draw, step, spin, stream).
// Stateless, so any spin's stream can be regenerated from its coordinates:
// results are the same whatever the launch shape, batch or step chunking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Host builds of this header (the CPU build of the walk core used by the
// tests) take portable equivalents of the device intrinsics below.
#if defined(__HIP_DEVICE_COMPILE__)
#define NZ_HD __host__ __device__
__device__ __forceinline__ uint32_t nz_umulhi(uint32_t a, uint32_t b) { return __umulhi(a, b); }
__device__ __forceinline__ void nz_sincospif(float x, float* s, float* c) { sincospif(x, s, c); }
__device__ __forceinline__ float nz_cospif(float x) { return cospif(x); }
__device__ __forceinline__ float nz_rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float nz_sqrtf(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ void nz_sincospi(double x, double* s, double* c) { sincospi(x, s, c); }
__device__ __forceinline__ double nz_cospi(double x) { return cospi(x); }
#else
#include <cmath>
#define NZ_HD __host__ __device__
inline uint32_t nz_umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline void nz_sincospif(float x, float* s, float* c) {
  *s = (float)std::sin(M_PI * (double)x);
  *c = (float)std::cos(M_PI * (double)x);
}
inline float nz_cospif(float x) { return (float)std::cos(M_PI * (double)x); }
inline float nz_rcpf(float x) { return 1.0f / x; }
inline float nz_sqrtf(float x) { return std::sqrt(x); }
inline void nz_sincospi(double x, double* s, double* c) {
  *s = std::sin(M_PI * x);
  *c = std::cos(M_PI * x);
}
inline double nz_cospi(double x) { return std::cos(M_PI * x); }
#endif

namespace nz {

enum : uint32_t { kStreamWalk = 0u, kStreamSeed = 1u };

struct u32x4 {
  uint32_t x, y, z, w;
};

NZ_HD inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = nz_umulhi(0xD2511F53u, c.x);
    const uint32_t lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = nz_umulhi(0xCD9E8D57u, c.z);
    const uint32_t lo1 = 0xCD9E8D57u * c.z;
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// ((x >> 9) + 1/2) * 2^-23: 23-bit uniform on (0,1), exact in binary32.
NZ_HD inline float unif23(uint32_t x) {
  return (float)(((x >> 9) << 1) | 1u) * 5.9604644775390625e-8f;  // * 2^-24
}

// Three standard normals (Box-Muller on (u0,u1) and (u2,u3)).
NZ_HD inline void normals3(uint64_t seed, uint32_t stream, uint32_t spin,
                                         uint32_t step, uint32_t draw, float& n0, float& n1,
                                         float& n2) {
  const u32x4 r = philox4x32_10(u32x4{draw, step, spin, stream}, (uint32_t)seed,
                                (uint32_t)(seed >> 32));
  const float r0 = sqrtf(-2.0f * logf(unif23(r.x)));
  const float r1 = sqrtf(-2.0f * logf(unif23(r.z)));
  float s0, c0;
  nz_sincospif(2.0f * unif23(r.y), &s0, &c0);
  n0 = r0 * c0;
  n1 = r0 * s0;
  n2 = r1 * nz_cospif(2.0f * unif23(r.w));
}

// The same three normals in double precision (what the walker uses).
NZ_HD inline void normals3_d(uint64_t seed, uint32_t stream, uint32_t spin, uint32_t step,
                             uint32_t draw, double& n0, double& n1, double& n2) {
  const u32x4 r = philox4x32_10(u32x4{draw, step, spin, stream}, (uint32_t)seed,
                                (uint32_t)(seed >> 32));
  const double r0 = sqrt(-2.0 * log((double)unif23(r.x)));
  const double r1 = sqrt(-2.0 * log((double)unif23(r.z)));
  double s0, c0;
  nz_sincospi(2.0 * (double)unif23(r.y), &s0, &c0);
  n0 = r0 * c0;
  n1 = r0 * s0;
  n2 = r1 * nz_cospi(2.0 * (double)unif23(r.w));
}

}  // namespace nz
