// The blocked phase contraction of many gradient waveforms (the precision
// form of the DWI walker).
//
// The reference accumulates, every time step, phase[w,i] += (x_i . u_w) G_w
// for each direction (randomWalk3d_phase_multidirr, sim3d_kernel.cu:880-887):
// W multiply-adds per spin-step on the CUDA cores.  With W different waveform
// shapes (not only directions of one shape) there is no 3-vector shortcut,
// but K steps can be blocked: with the displacements d_i(t) of K steps as a
// row of 3K values,
//     phase[i, w] += sum_{t,a} d_i,a(t) * g_w(t) u_w,a,
// an [N x 3K] by [3K x W] product.  It runs here on the MFMA units as split
// BF16: x = hi + lo with hi = bf16(x), lo = bf16(x - hi) for both operands,
// and x.y ~ hi.hi + hi.lo + lo.hi accumulated in FP32 (relative error ~2^-16
// per product).  The start position enters once, phase[i, w] = (x0_i . u_w)
// sum_t g_w(t) before the walk, so the contraction only sees displacements
// (tens of um at most) and no wrap corrections are needed: d is unwrapped.
// An FP32 VALU contraction of the same blocks is kept for comparison.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nz_internal.h"

namespace nz {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 bits, and back
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// Waveform block: Bt[w][k] (k = 3 (t - t0) + a, K3 = 3K contiguous per
// waveform) = g[w][t] * u[a][w], split into hi / lo BF16.  Also the FP32
// values for the VALU contraction.
__global__ void __launch_bounds__(256) waveform_block_kernel(const float* g, int64_t T,
                                                             const float* u, int32_t W, int64_t t0,
                                                             int32_t K3, uint16_t* bt_hi,
                                                             uint16_t* bt_lo, float* bt_f32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)W * K3) return;
  const int w = (int)(e / K3), k = (int)(e % K3);
  const int64_t t = t0 + k / 3;
  const int a = k % 3;
  const float v = (t < T) ? g[(int64_t)w * T + t] * u[(int64_t)a * W + w] : 0.0f;
  const uint16_t h = bf16_bits(v);
  bt_hi[e] = h;
  bt_lo[e] = bf16_bits(v - bf16_float(h));
  if (bt_f32) bt_f32[e] = v;
}

constexpr int kTileS = 64;   // slots per workgroup tile
constexpr int kTileW = 64;   // waveforms per workgroup tile
constexpr int kChunk = 32;   // k per MFMA (16x16x32 BF16)
constexpr int kMaxK3 = 384;  // K <= 128: A operands held in registers

// phase[spin(s), w] += sum_k rec[s][k] * Bt[w][k] over one block of K3 =
// 32 NKC rows, 16x16x32 BF16 MFMAs, split hi/lo (three products).
// Workgroup: 4 waves and 64 slots; wave q owns slots [16q, 16q + 16) and
// walks every waveform in tiles of 64.
// - A (the block's displacements) is read from HBM once, straight into the
//   registers that hold it as MFMA A operands for all waveform tiles (lane l:
//   slot l % 16 of the wave, k = 8 (l / 16) .. + 8 of each 32-k chunk: 32
//   contiguous bytes of the slot's record, the 4 lanes of a slot 128).
// - B (the waveform block, a few hundred KB shared by all workgroups, served
//   from L2) moves through two LDS buffers, one 32-k chunk of one waveform
//   tile at a time: the next chunk's global load is in flight while the
//   current chunk's MFMAs run, and one barrier per chunk suffices.
// - D goes through LDS so each phase row is read-modified-written by 64 lanes
//   at once (256 contiguous bytes); each wave touches only its own 16 rows.
// 33 KB of LDS and 168 VGPRs: three workgroups per CU.
// MFMA operand layout (lane l of 64): A row l % 16, k 8 (l / 16) .. + 8;
// B column l % 16, the same k; D column l % 16, rows 4 (l / 16) .. + 4.
// Position of k (a multiple of 8) in LDS row `col` of a waveform chunk: rows
// of 32 BF16 (64 bytes, four 16-byte pieces) with the piece index XORed with
// bits 1-2 of the row, so that the MFMA B reads (16 rows, one piece each) and
// the chunk stores (4 pieces of 16 rows) fall in distinct LDS banks (the
// bank-conflict counter reads ~0 against 44% of LDS cycles with padded rows).
__device__ __forceinline__ int sbk(int col, int k) { return 8 * ((k >> 3) ^ ((col >> 1) & 3)); }

template <int NKC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
phase_gemm_mfma_kernel(const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin,
                       const uint16_t* bt_hi, const uint16_t* bt_lo, int32_t W, float* phase) {
  constexpr int K3 = NKC * kChunk;
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][2][kTileW][kChunk];  // [buf][hi, lo]
  __shared__ float sC[kTileS][kTileW + 1];
  const int64_t s0 = (int64_t)blockIdx.x * kTileS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kl = 8 * (lane >> 4);
  const int64_t srow = s0 + wave * 16 + (lane & 15);
  const bool row_ok = srow < nslots;
  bf16x8 ah[NKC], al[NKC];
#pragma unroll
  for (int kc = 0; kc < NKC; ++kc) {
    uint16_t hb[8], lb[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float v = row_ok ? rec[srow * ld + kc * kChunk + kl + j] : 0.0f;
      hb[j] = bf16_bits(v);
      lb[j] = bf16_bits(v - bf16_float(hb[j]));
    }
    __builtin_memcpy(&ah[kc], hb, 16);
    __builtin_memcpy(&al[kc], lb, 16);
  }
  // the wave's 16 slots -> spins
  int32_t spin_row = -1;
  if (lane < 16) {
    const int64_t s = s0 + wave * 16 + lane;
    spin_row = (s < nslots) ? slot_spin[s] : -1;
  }
  // this thread's 16-byte piece of every B chunk: waveform ww, k q .. q + 8
  const int ww = tid >> 2, q = (tid & 3) * 8;
  bf16x8 nh = {}, nl = {};  // the next chunk, in flight
  if (ww < W) {
    __builtin_memcpy(&nh, &bt_hi[(int64_t)ww * K3 + q], 16);
    __builtin_memcpy(&nl, &bt_lo[(int64_t)ww * K3 + q], 16);
  }
  int buf = 0;
  for (int w0 = 0; w0 < W; w0 += kTileW) {
    floatx4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      // buffer `buf` was last read two chunks ago, before the previous barrier
      __builtin_memcpy(&sB[buf][0][ww][sbk(ww, q)], &nh, 16);
      __builtin_memcpy(&sB[buf][1][ww][sbk(ww, q)], &nl, 16);
      {
        const int nw0 = (kc + 1 < NKC) ? w0 : w0 + kTileW;
        const int nkc = (kc + 1 < NKC) ? kc + 1 : 0;
        if (nw0 < W) {
          nh = bf16x8{};
          nl = bf16x8{};
          if (nw0 + ww < W) {
            const int64_t o = (int64_t)(nw0 + ww) * K3 + nkc * kChunk + q;
            __builtin_memcpy(&nh, &bt_hi[o], 16);
            __builtin_memcpy(&nl, &bt_lo[o], 16);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int col = c * 16 + (lane & 15);
        bf16x8 bh, bl;
        __builtin_memcpy(&bh, &sB[buf][0][col][sbk(col, kl)], 16);
        __builtin_memcpy(&bl, &sB[buf][1][col][sbk(col, kl)], 16);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[kc], bh, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[kc], bl, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[kc], bh, acc[c], 0, 0, 0);
      }
      buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < 4; ++c) sC[wave * 16 + 4 * (lane >> 4) + i][c * 16 + (lane & 15)] = acc[c][i];
    __syncthreads();
    const int w = w0 + lane;
    float old[16];  // all 16 row loads in flight before the first add
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      const int32_t sp = __shfl(spin_row, rr, 64);
      old[rr] = (sp >= 0 && w < W) ? phase[(int64_t)sp * W + w] : 0.0f;
    }
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      const int32_t sp = __shfl(spin_row, rr, 64);
      if (sp >= 0 && w < W) phase[(int64_t)sp * W + w] = old[rr] + sC[wave * 16 + rr][lane];
    }
  }
}

// The same contraction on the vector ALUs in FP32 (the CUDA-core form).
// Thread: one slot, 16 waveforms; the block's waveform columns via LDS.
__global__ void __launch_bounds__(256) phase_gemm_fp32_kernel(
    const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin, int32_t K3,
    const float* bt, int32_t W, float* phase) {
  __shared__ float sB[16][kChunk];
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int w0 = blockIdx.y * 16;
  float acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.0f;
  for (int k0 = 0; k0 < K3; k0 += kChunk) {
    for (int e = threadIdx.x; e < 16 * kChunk; e += blockDim.x) {
      const int ww = e / kChunk, kk = e % kChunk;
      const int w = w0 + ww;
      sB[ww][kk] = (w < W && k0 + kk < K3) ? bt[(int64_t)w * K3 + k0 + kk] : 0.0f;
    }
    __syncthreads();
    if (s < nslots) {
      for (int kk = 0; kk < kChunk && k0 + kk < K3; ++kk) {
        const float x = rec[s * ld + k0 + kk];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] = fmaf(x, sB[c][kk], acc[c]);
      }
    }
    __syncthreads();
  }
  if (s >= nslots) return;
  const int32_t spin = slot_spin[s];
  if (spin < 0) return;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const int w = w0 + c;
    if (w < W) phase[(int64_t)spin * W + w] += acc[c];
  }
}

// phase[i, w] = (x0_i . u_w) * sum_t g_w(t): the start position's share.
// Grid: x over spins in groups of kInitSpins, y over waveforms in tiles of
// 256 (one thread per waveform, so each spin's row is written coalesced);
// no dimension grows with n * W.
constexpr int kInitSpins = 16;
__global__ void __launch_bounds__(256) phase_init_kernel(const float* x0, int64_t n, const float* u,
                                                         const float* gsum, int32_t W,
                                                         float* phase) {
  const int w = blockIdx.y * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const float ux = u[w], uy = u[W + w], uz = u[2 * W + w], gs = gsum[w];
  const int64_t i0 = (int64_t)blockIdx.x * kInitSpins;
  for (int64_t i = i0; i < i0 + kInitSpins && i < n; ++i)
    phase[i * W + w] = (x0[i] * ux + x0[n + i] * uy + x0[2 * n + i] * uz) * gs;
}

// out[2w], out[2w+1] += sum over a block of spins of cos / sin phase[i, w]
// (FP32 partial sums of 32 spins, then FP64).
__global__ void __launch_bounds__(256) phase_signal_kernel(const float* phase, int64_t n, int32_t W,
                                                           int64_t per, double* out) {
  const int w = blockIdx.y * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = (i0 + per < n) ? i0 + per : n;
  double sc = 0.0, ss = 0.0;
  float pc = 0.0f, ps = 0.0f;
  int cnt = 0;
  for (int64_t i = i0; i < i1; ++i) {
    float s, c;
    sincosf(phase[i * W + w], &s, &c);
    pc += c;
    ps += s;
    if (++cnt == 32) {
      sc += pc;
      ss += ps;
      pc = ps = 0.0f;
      cnt = 0;
    }
  }
  sc += pc;
  ss += ps;
  atomicAdd(&out[2 * w], sc);
  atomicAdd(&out[2 * w + 1], ss);
}

__global__ void zero_f64_phase(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace nz

using namespace nz;

extern "C" {

int nz_waveform_block(const float* g, int64_t T, const float* u, int32_t W, int64_t t0, int32_t K,
                      uint16_t* bt_hi, uint16_t* bt_lo, float* bt_f32, void* stream) {
  if (!g || !u || !bt_hi || !bt_lo) return fail(-1, "nz_waveform_block: NULL buffer");
  if (T <= 0 || W <= 0 || K <= 0 || t0 < 0) return fail(-1, "nz_waveform_block: bad sizes");
  const int32_t K3 = 3 * K;
  const int64_t total = (int64_t)W * K3;
  hipLaunchKernelGGL(waveform_block_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, g, T, u, W, t0, K3, bt_hi, bt_lo, bt_f32);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_phase_gemm(const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin,
                  int32_t K, const uint16_t* bt_hi, const uint16_t* bt_lo, const float* bt_f32,
                  int32_t W, int32_t mode, float* phase, void* stream) {
  if (!rec || !slot_spin || !phase) return fail(-1, "nz_phase_gemm: NULL buffer");
  const int32_t K3 = 3 * K;
  if (nslots <= 0 || ld < K3 || W <= 0 || K <= 0) return fail(-1, "nz_phase_gemm: bad sizes");
  if (mode == 0) {
    if (!bt_hi || !bt_lo) return fail(-1, "nz_phase_gemm: MFMA mode needs bt_hi and bt_lo");
    if (K3 % kChunk || K3 > kMaxK3)
      return fail(-1, "nz_phase_gemm: MFMA mode needs K a multiple of 32 and at most 128");
    dim3 grid((unsigned)((nslots + kTileS - 1) / kTileS));
    auto* kern = (K3 == 96)    ? phase_gemm_mfma_kernel<3>
                 : (K3 == 192) ? phase_gemm_mfma_kernel<6>
                 : (K3 == 288) ? phase_gemm_mfma_kernel<9>
                               : phase_gemm_mfma_kernel<12>;
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, (hipStream_t)stream, rec, ld, nslots, slot_spin,
                       bt_hi, bt_lo, W, phase);
  } else {
    if (!bt_f32) return fail(-1, "nz_phase_gemm: FP32 mode needs bt_f32");
    dim3 grid((unsigned)((nslots + 255) / 256), (unsigned)((W + 15) / 16));
    hipLaunchKernelGGL(phase_gemm_fp32_kernel, grid, dim3(256), 0, (hipStream_t)stream, rec, ld,
                       nslots, slot_spin, K3, bt_f32, W, phase);
  }
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_phase_init(const float* x0, int64_t n, const float* u, const float* gsum, int32_t W,
                  float* phase, void* stream) {
  if (!x0 || !u || !gsum || !phase || n <= 0 || W <= 0) return fail(-1, "nz_phase_init: bad arguments");
  const int64_t gx = (n + kInitSpins - 1) / kInitSpins, gy = ((int64_t)W + 255) / 256;
  if (gx > (int64_t)INT32_MAX || gy > 65535)
    return fail(-1, "nz_phase_init: n or W too large for one launch");
  hipLaunchKernelGGL(phase_init_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0,
                     (hipStream_t)stream, x0, n, u, gsum, W, phase);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_phase_signals(const float* phase, int64_t n, int32_t W, double* out, void* stream) {
  if (!phase || !out || n <= 0 || W <= 0) return fail(-1, "nz_phase_signals: bad arguments");
  hipLaunchKernelGGL(zero_f64_phase, dim3((unsigned)((2 * W + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, out, (int64_t)2 * W);
  const int64_t per = 4096;
  dim3 grid((unsigned)((n + per - 1) / per), (unsigned)((W + 255) / 256));
  hipLaunchKernelGGL(phase_signal_kernel, grid, dim3(256), 0, (hipStream_t)stream, phase, n, W,
                     per, out);
  NZ_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
