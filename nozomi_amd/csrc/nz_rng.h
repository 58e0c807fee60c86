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
__device__ __forceinline__ float nz_rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float nz_sqrtf(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
#include <cmath>
#define NZ_HD __host__ __device__
inline float nz_rcpf(float x) { return 1.0f / x; }
inline float nz_sqrtf(float x) { return std::sqrt(x); }
#endif

namespace nz {

enum : uint32_t { kStreamWalk = 0u, kStreamSeed = 1u };

struct u32x4 {
  uint32_t x, y, z, w;
};

// Both halves of the 64-bit product as one multiply (v_mad_u64_u32 on the
// device rather than separate high and low multiplies).
NZ_HD inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

NZ_HD inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c.x, hi0, lo0);
    mulhilo(0xCD9E8D57u, c.z, hi1, lo1);
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

// Box-Muller in binary32 with the same bits on every IEEE-754 machine: only
// correctly rounded operations (+ - * /, fma, sqrt), evaluated without
// contraction, and polynomials written out, so the device and a host build
// of this file (and the test oracle, which restates it) agree exactly.

// ln u for u in [2^-24, 1): u = m 2^e with m in [1/sqrt2, sqrt2); m - 1 is
// exact (Sterbenz); ln m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172, by
// its Taylor series to s^9 (truncation < 4e-10); ln 2 split in a 17-bit head
// (e * head exact) and a tail.
NZ_HD inline float nz_ln_u(float u) {
#pragma clang fp contract(off)
  uint32_t b;
  __builtin_memcpy(&b, &u, 4);
  int e = (int)((b >> 23) & 0xffu) - 127;
  const uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
  float m;
  __builtin_memcpy(&m, &mb, 4);  // [1, 2)
  if (m > 1.41421354f) {
    m = m * 0.5f;  // exact
    e += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float z = s * s;
  float p = fmaf(z, 0.111111112f, 0.142857149f);  // 1/9, 1/7
  p = fmaf(z, p, 0.200000003f);                    // 1/5
  p = fmaf(z, p, 0.333333343f);                    // 1/3
  p = fmaf(z, p, 1.0f);
  const float lm = (2.0f * s) * p;
  const float fe = (float)e;
  return fmaf(fe, 0.693145752f, fmaf(fe, 1.42860677e-06f, lm));
}

// sin and cos of 2 pi v for v in (0, 1) on the 2^-24 grid: v = q/4 + a with
// q = round(4v) and |a| <= 1/8 turn, both exact; Taylor polynomials of
// sin / cos(2 pi a) to degree 9 / 10 (|2 pi a| <= pi/4: truncation < 3e-9),
// then the quadrant rotation.
NZ_HD inline void nz_sincos_2pi(float v, float* sn, float* cs) {
#pragma clang fp contract(off)
  const float q = floorf(v * 4.0f + 0.5f);
  const float a = v - q * 0.25f;
  const float z = a * a;
  float ps = fmaf(z, 42.0586929f, -76.7058563f);
  ps = fmaf(z, ps, 81.6052475f);
  ps = fmaf(z, ps, -41.3417015f);
  ps = fmaf(z, ps, 6.28318548f);
  const float sa = a * ps;
  float pc = fmaf(z, -26.4262562f, 60.2446404f);
  pc = fmaf(z, pc, -85.4568176f);
  pc = fmaf(z, pc, 64.9393921f);
  pc = fmaf(z, pc, -19.7392082f);
  const float ca = fmaf(z, pc, 1.0f);
  switch (((int)q) & 3) {
    case 0: *sn = sa; *cs = ca; break;
    case 1: *sn = ca; *cs = -sa; break;
    case 2: *sn = -sa; *cs = -ca; break;
    default: *sn = -ca; *cs = sa; break;
  }
}

// Three standard normals for one (seed, stream, spin, step, draw) counter:
// Box-Muller on (u0, u1) gives n0, n1; on (u2, u3) gives n2.
NZ_HD inline void normals3(uint64_t seed, uint32_t stream, uint32_t spin, uint32_t step,
                           uint32_t draw, float& n0, float& n1, float& n2) {
#pragma clang fp contract(off)
  const u32x4 r = philox4x32_10(u32x4{draw, step, spin, stream}, (uint32_t)seed,
                                (uint32_t)(seed >> 32));
  const float r0 = sqrtf(-2.0f * nz_ln_u(unif23(r.x)));
  const float r1 = sqrtf(-2.0f * nz_ln_u(unif23(r.z)));
  float s0, c0, s1, c1;
  nz_sincos_2pi(unif23(r.y), &s0, &c0);
  nz_sincos_2pi(unif23(r.w), &s1, &c1);
  n0 = r0 * c0;
  n1 = r0 * s0;
  n2 = r1 * c1;
}

}  // namespace nz
