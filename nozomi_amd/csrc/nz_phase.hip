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

// phase[spin(s), w] += sum_k rec[k][s] * Bt[w][k] over one block of K3 rows,
// 16x16x32 BF16 MFMAs, split hi/lo (three products).  Workgroup: 4 waves and
// 64 slots; wave q owns slots [16q, 16q + 16) and walks every waveform in
// tiles of 64.  The block's displacements are read from HBM once: staged
// through LDS (coalesced rows -> per-lane k-contiguous vectors), split, and
// kept in registers as the A operands of all waveform tiles; the B operands
// (the waveform block, a few hundred KB shared by all workgroups) come from L2.
// MFMA operand layout (lane l of 64): A row l % 16, k 8 (l / 16) .. + 8;
// B column l % 16, the same k; D column l % 16, rows 4 (l / 16) .. + 4.
constexpr int kStageK3 = 192;  // rows of the block staged through LDS at a time
constexpr int kMaxK3 = 384;    // K <= 128: A operands held in registers

__global__ void __launch_bounds__(256) phase_gemm_mfma_kernel(
    const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin, int32_t K3,
    const uint16_t* bt_hi, const uint16_t* bt_lo, int32_t W, float* phase) {
  __shared__ __attribute__((aligned(16))) uint16_t sAh[kTileS][kStageK3 + 8];
  __shared__ __attribute__((aligned(16))) uint16_t sAl[kTileS][kStageK3 + 8];
  // one k-chunk of the waveform block, shared by the four waves
  __shared__ __attribute__((aligned(16))) uint16_t sBh[kTileW][kChunk + 8];
  __shared__ __attribute__((aligned(16))) uint16_t sBl[kTileW][kChunk + 8];
  const int64_t s0 = (int64_t)blockIdx.x * kTileS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nkc = K3 / kChunk;
  const int r = wave * 16 + (lane & 15);
  const int kl = 8 * (lane >> 4);
  bf16x8 ah[kMaxK3 / kChunk], al[kMaxK3 / kChunk];
#pragma unroll
  for (int st = 0; st < kMaxK3 / kStageK3; ++st) {  // stage up to 192 rows at a time
    const int kh = st * kStageK3;  // (unrolled: register indices stay static)
    if (kh < K3) {
      const int rows = (K3 - kh < kStageK3) ? K3 - kh : kStageK3;
      for (int e = tid; e < rows * kTileS; e += 256) {
        const int kk = e / kTileS, ss = e % kTileS;
        const int64_t s = s0 + ss;
        const float v = (s < nslots) ? rec[(int64_t)(kh + kk) * ld + s] : 0.0f;
        const uint16_t h = bf16_bits(v);
        sAh[ss][kk] = h;
        sAl[ss][kk] = bf16_bits(v - bf16_float(h));
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kStageK3 / kChunk; ++j) {
        const int kc = st * (kStageK3 / kChunk) + j;
        if (kc < nkc) {
          __builtin_memcpy(&ah[kc], &sAh[r][j * kChunk + kl], 16);
          __builtin_memcpy(&al[kc], &sAl[r][j * kChunk + kl], 16);
        }
      }
      __syncthreads();
    }
  }
  // the wave's 16 slots -> spins (row r of the tile: slot s0 + r)
  int32_t spin_row = -1;
  if (lane < 16) {
    const int64_t s = s0 + wave * 16 + lane;
    spin_row = (s < nslots) ? slot_spin[s] : -1;
  }
  float(*sC)[kTileW + 1] = reinterpret_cast<float(*)[kTileW + 1]>(&sAh[0][0]);
  for (int w0 = 0; w0 < W; w0 += kTileW) {
    floatx4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kc = 0; kc < kMaxK3 / kChunk; ++kc) {
      if (kc < nkc) {
      {  // 64 waveforms x 32 k, hi and lo: 16 bytes per thread each
        const int ww = tid >> 2, q = (tid & 3) * 8;
        const int w = w0 + ww;
        bf16x8 vh = {}, vl = {};
        if (w < W) {
          const int64_t o = (int64_t)w * K3 + kc * kChunk + q;
          __builtin_memcpy(&vh, &bt_hi[o], 16);
          __builtin_memcpy(&vl, &bt_lo[o], 16);
        }
        __builtin_memcpy(&sBh[ww][q], &vh, 16);
        __builtin_memcpy(&sBl[ww][q], &vl, 16);
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int col = c * 16 + (lane & 15);
        bf16x8 bh, bl;
        __builtin_memcpy(&bh, &sBh[col][kl], 16);
        __builtin_memcpy(&bl, &sBl[col][kl], 16);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[kc], bh, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[kc], bl, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[kc], bh, acc[c], 0, 0, 0);
      }
      __syncthreads();
      }
    }
    // D (column lane % 16 of each 16x16 tile, rows 4 (lane / 16) + i) goes
    // through LDS so that each phase row is updated by 64 lanes at once:
    // 256 contiguous bytes per read-modify-write instead of 16 scattered
    // 64-byte pieces.  Each wave only touches its own 16 rows.
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
    __syncthreads();
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
        const float x = rec[(int64_t)(k0 + kk) * ld + s];
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
__global__ void __launch_bounds__(256) phase_init_kernel(const float* x0, int64_t n, const float* u,
                                                         const float* gsum, int32_t W,
                                                         float* phase) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * W) return;
  const int64_t i = e / W;
  const int w = (int)(e % W);
  phase[e] = (x0[i] * u[w] + x0[n + i] * u[W + w] + x0[2 * n + i] * u[2 * W + w]) * gsum[w];
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
  if (nslots <= 0 || ld < nslots || W <= 0 || K <= 0) return fail(-1, "nz_phase_gemm: bad sizes");
  const int32_t K3 = 3 * K;
  if (mode == 0) {
    if (!bt_hi || !bt_lo) return fail(-1, "nz_phase_gemm: MFMA mode needs bt_hi and bt_lo");
    if (K3 % kChunk || K3 > kMaxK3)
      return fail(-1, "nz_phase_gemm: MFMA mode needs K a multiple of 32 and at most 128");
    dim3 grid((unsigned)((nslots + kTileS - 1) / kTileS));
    hipLaunchKernelGGL(phase_gemm_mfma_kernel, grid, dim3(256), 0, (hipStream_t)stream, rec, ld,
                       nslots, slot_spin, K3, bt_hi, bt_lo, W, phase);
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
  const int64_t total = n * W;
  hipLaunchKernelGGL(phase_init_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
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
