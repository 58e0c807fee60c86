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

typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---- split-BF16 contraction on 32x32x16 MFMAs, wave-specialised ----------
//
// phase[spin(s), w] += sum_k rec[s][k] * Bt[w][k] over one block of K3 = 3K
// values per slot.  Workgroup: 128 slots (four groups of 32) and three kinds
// of wave:
//  - 4 KS MFMA waves: wave (g, q) holds the A operand (slots 32g .. 32g+31,
//    k in part q of KS equal parts, split hi/lo) in registers for the whole
//    kernel, and accumulates a 32 x 32 tile of (slot, waveform) sums with
//    v_mfma_f32_32x32x16_bf16: hi.hi + hi.lo + lo.hi per 16-k step;
//  - one waveform wave streams the waveform block (L2-resident) into an LDS
//    ring of kRing entries with global_load_lds, entry = the 32-k chunk of
//    every part for one waveform tile, issued kRing - 1 chunks ahead;
//  - one phase wave streams the old phases of each tile (32 waveforms x 128
//    rows, gathered by spin) into LDS one tile ahead.
// Each wave's loads retire in issue order (vmcnt), so HBM latency (the phase
// rows) and L2 latency (the waveform chunks) sit in different waves, and the
// MFMA waves issue no loads after their A operand.  After a tile's chunks the
// parts' tile sums wait in an LDS stage; during the next tile's first two
// chunk intervals they are added into the old phases in a fixed order (part
// 0 first) and the rows stored back, 128 contiguous bytes per row.  Barrier
// per chunk interval (s_barrier, every wave the same count).
// Measured at 4e6 slots x 1024 waveforms: v_mfma_f32_32x32x16_bf16 sustains
// ~2.0 PFLOP/s against 1.2 for v_mfma_f32_16x16x32_bf16, and the 32-row tile
// halves the LDS reads per MFMA.
// MFMA operand layout (lane l of 64): A row l % 32, k 8 (l / 32) .. + 8;
// B column l % 32, the same k; D column l % 32, rows 8 (v / 4) + 4 (l / 32)
// + v % 4 for v = 0..15.
constexpr int kTileS = 128;  // slots per workgroup
constexpr int kTileW = 32;   // waveforms per tile

// Position of k (a multiple of 8) in LDS row `col` of a waveform chunk of CH
// k (32 or 64): rows of CH BF16 (CH / 8 16-byte pieces), the piece index
// XORed with bits 2-3 of the row (CH = 32) or with bits 0-2 of the row and
// bit 3 (CH = 64), so the rows of any 16 consecutive lanes of an MFMA B read
// fall on 16 different 16-byte bank groups and the 1 KB read uses every bank
// the minimum number of times; the ring stores (1 KB of whole rows each)
// likewise.
template <int CH>
__device__ __forceinline__ int bswz(int col) {
  return CH == 32 ? ((col >> 2) & 3) : ((col & 7) ^ ((col >> 3) & 1));
}
// Word of (row r, column c) of a 128 x 32 FP32 phase tile in LDS: rows of 32
// words, the 4-word group index XORed with 4 ((r / 4) % 2).  Phase-wave load
// j (0..15) of lane l covers words 256 j + 4 l .. + 4: row 8 j + l / 8,
// columns 4 ((l % 8) ^ 4 ((l / 32) % 2)) .. + 4, and the MFMA waves store
// the same pieces back.
__device__ __forceinline__ int pword(int r, int c) {
  return r * kTileW + ((((c >> 2) ^ (((r >> 2) & 1) << 2))) << 2) + (c & 3);
}
__device__ __forceinline__ int ppiece_col(int lane) { return ((lane & 7) ^ (4 * ((lane >> 5) & 1))) << 2; }

// Wait until at most n (a multiple of 4, <= 60) vector memory operations of
// this wave are outstanding.
__device__ __forceinline__ void wait_vmcnt(int n) {
  switch (n >> 2) {
    case 0: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    case 1: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
    case 2: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
    case 3: asm volatile("s_waitcnt vmcnt(12)" ::: "memory"); break;
    case 4: asm volatile("s_waitcnt vmcnt(16)" ::: "memory"); break;
    case 5: asm volatile("s_waitcnt vmcnt(20)" ::: "memory"); break;
    case 6: asm volatile("s_waitcnt vmcnt(24)" ::: "memory"); break;
    case 7: asm volatile("s_waitcnt vmcnt(28)" ::: "memory"); break;
    case 8: asm volatile("s_waitcnt vmcnt(32)" ::: "memory"); break;
    case 9: asm volatile("s_waitcnt vmcnt(36)" ::: "memory"); break;
    case 10: asm volatile("s_waitcnt vmcnt(40)" ::: "memory"); break;
    case 11: asm volatile("s_waitcnt vmcnt(44)" ::: "memory"); break;
    case 12: asm volatile("s_waitcnt vmcnt(48)" ::: "memory"); break;
    case 13: asm volatile("s_waitcnt vmcnt(52)" ::: "memory"); break;
    case 14: asm volatile("s_waitcnt vmcnt(56)" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(60)" ::: "memory"); break;
  }
}
// LDS writes done, then the workgroup barrier.
__device__ __forceinline__ void wg_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// NKC chunk intervals of CH k per tile, KS parts of K: K3 = CH NKC KS.  A
// chunk interval ends with a workgroup barrier: 64-k chunks halve the
// barriers per tile (the MFMA waves read a chunk's B operands in two halves
// into the same registers), with a ring of four entries instead of six.
template <int NKC, int KS, int CH>
__global__ void __launch_bounds__(64 * (4 * KS + 2))
__attribute__((amdgpu_waves_per_eu(KS == 3 ? 4 : 3)))
phase_gemm_mfma_kernel(const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin,
                       const uint16_t* bt_hi, const uint16_t* bt_lo, int32_t W, float* phase) {
  constexpr int kChunk = CH;
  static_assert(CH == 32 || CH == 64, "chunks of 32 or 64 k");
  constexpr int kRing = CH == 32 ? 6 : 4;  // waveform chunk entries in LDS
  constexpr int KP = NKC * kChunk;  // k per part
  constexpr int K3 = KP * KS;
  constexpr int kEntry = KS * 2 * kTileW * kChunk;  // uint16 per ring entry
  constexpr int kEntryLoads = KS * 2 * (kChunk / 16);  // 1 KB loads per ring entry
  constexpr int kTileWords = kTileS * kTileW;
  // Each part writes its tile sums to an LDS stage at the end of the tile,
  // and in the first two chunk intervals of the next tile every MFMA wave
  // adds old + part 0 + part 1 ... for its share of the rows, half of it per
  // interval, and stores them, beside that tile's MFMAs (a per-part
  // epilogue, each part adding its tile into the old phases in LDS behind a
  // barrier of its own, ran K = 64 8.7, K = 96 11.9, K = 128 12.8 ms per
  // block against 7.4, 10.4, 11.1 ms; the same additions in the same order,
  // so the same bits).
  static_assert(NKC >= 3, "the staged epilogue needs three chunk intervals per tile");
  __shared__ __attribute__((aligned(16))) uint16_t sB[kRing * kEntry];
  __shared__ __attribute__((aligned(16))) float sP[2 * kTileWords];
  __shared__ __attribute__((aligned(16))) float sS[KS * kTileWords];
  const int64_t s0 = (int64_t)blockIdx.x * kTileS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ntiles = (W + kTileW - 1) / kTileW;
  const int nint = ntiles * NKC;  // chunk intervals

  if (wave < 4 * KS) {  // ---- MFMA wave (slot group sg, part kq)
    const int sg = wave & 3, kq = wave >> 2;
    const int kh = 8 * (lane >> 5), col = lane & 31;
    bf16x8 ah[NKC * (kChunk / 16)], al[NKC * (kChunk / 16)];
    {
      const int64_t srow = s0 + sg * 32 + col;
      const bool row_ok = srow < nslots;
      const float* arow = rec + (row_ok ? srow : 0) * ld + kq * KP + kh;
#pragma unroll
      for (int ks = 0; ks < NKC * (kChunk / 16); ++ks) {
        floatx4 v[2];
        __builtin_memcpy(&v[0], arow + ks * 16, 16);
        __builtin_memcpy(&v[1], arow + ks * 16 + 4, 16);
        uint16_t hb[8], lb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const float x = row_ok ? v[j >> 2][j & 3] : 0.0f;
          hb[j] = bf16_bits(x);
          lb[j] = bf16_bits(x - bf16_float(hb[j]));
        }
        __builtin_memcpy(&ah[ks], hb, 16);
        __builtin_memcpy(&al[ks], lb, 16);
      }
    }
    int32_t spin_row;
    {
      const int64_t s = s0 + sg * 32 + col;
      const int32_t v = slot_spin[s < nslots ? s : nslots - 1];
      spin_row = (s < nslots) ? v : -1;
    }
    int32_t psp[4];  // spins of this lane's four store pieces (rows 32 sg + 8 j + l / 8)
#pragma unroll
    for (int j = 0; j < 4; ++j) psp[j] = __shfl(spin_row, 8 * j + (lane >> 3), 64);
    const int pcol = ppiece_col(lane);
    int bo[kChunk / 16];  // the lane's B piece of 16-k step s: k = 16 s + kh of row col
#pragma unroll
    for (int s = 0; s < kChunk / 16; ++s) bo[s] = col * kChunk + 8 * ((2 * s + (kh >> 3)) ^ bswz<CH>(col));
    // old phases + the staged part sums of tile u, rows
    // 32 sg + 8 j + l / 8 for this wave's pieces j (half 0: the first half of
    // them, 1: the rest, 2: all), stored back
    auto combine = [&](int u, int half) {
      const float* tp = sP + (u & 1) * kTileWords;
      const int lo = (4 * kq) / KS, hi = (4 * (kq + 1)) / KS, mid = lo + (hi - lo + 1) / 2;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (half == 0 ? (j < lo || j >= mid) : half == 1 ? (j < mid || j >= hi) : (j < lo || j >= hi)) continue;
        const int off = (sg * 32 + 8 * j) * kTileW + 4 * lane;
        floatx4 v;
        __builtin_memcpy(&v, tp + off, 16);
#pragma unroll
        for (int q = 0; q < KS; ++q) {
          floatx4 a;
          __builtin_memcpy(&a, sS + q * kTileWords + off, 16);
          v += a;
        }
        const int c = u * kTileW + pcol;
        if (psp[j] >= 0 && c < W) __builtin_nontemporal_store(v, (floatx4*)(phase + (int64_t)psp[j] * W + c));
      }
    };
    wg_barrier();
    int g = 0;
    for (int t = 0; t < ntiles; ++t) {
      floatx16 acc;
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
#pragma unroll
      for (int kc = 0; kc < NKC; ++kc, ++g) {
        const uint16_t* bh = sB + (g % kRing) * kEntry + kq * (2 * kTileW * kChunk);
        const uint16_t* bl = bh + kTileW * kChunk;
#pragma unroll
        for (int s2 = 0; s2 < kChunk / 32; ++s2) {  // two 16-k steps per half
          const int a0 = kc * (kChunk / 16) + 2 * s2;
          bf16x8 h0, l0, h1, l1;
          __builtin_memcpy(&h0, bh + bo[2 * s2], 16);
          __builtin_memcpy(&l0, bl + bo[2 * s2], 16);
          __builtin_memcpy(&h1, bh + bo[2 * s2 + 1], 16);
          __builtin_memcpy(&l1, bl + bo[2 * s2 + 1], 16);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a0], h0, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a0], l0, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a0], h0, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a0 + 1], h1, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a0 + 1], l1, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a0 + 1], h1, acc, 0, 0, 0);
        }
        if (kc < 2 && t >= 1) combine(t - 1, kc);
        if (kc == NKC - 1) {
#pragma unroll
          for (int v = 0; v < 16; ++v)
            sS[kq * kTileWords + pword(sg * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), col)] = acc[v];
        }
        wg_barrier();
      }
    }
    combine(ntiles - 1, 2);
  } else if (wave == 4 * KS) {  // ---- waveform wave: ring entries
    // entry of interval gi: chunk gi % NKC of tile gi / NKC, all KS parts
    auto issue = [&](int gi) {
      const int t = gi / NKC, kc = gi - t * NKC;
      uint16_t* dst = sB + (gi % kRing) * kEntry;
#pragma unroll
      for (int q = 0; q < KS; ++q)
#pragma unroll
        for (int hl = 0; hl < 2; ++hl)
#pragma unroll
          for (int cg = 0; cg < kChunk / 16; ++cg) {  // 1 KB (64 pieces) per load
            constexpr int kPieces = kChunk / 8;          // 16-byte pieces per row
            const int e = cg * 64 + lane, c = e / kPieces;
            // LDS piece e % kPieces of row c holds k-piece kp
            const int kp = (e - c * kPieces) ^ bswz<CH>(c);
            const int w = t * kTileW + c;
            const uint16_t* src = (hl ? bt_lo : bt_hi) + (int64_t)(w < W ? w : 0) * K3 + q * KP +
                                  kc * kChunk + 8 * kp;
            __builtin_amdgcn_global_load_lds(src, dst + (q * 2 + hl) * (kTileW * kChunk) + cg * 512,
                                             16, 0, 0);
          }
    };
    const int first = nint < kRing ? nint : kRing;
    for (int gi = 0; gi < first; ++gi) issue(gi);
    wait_vmcnt(kEntryLoads * (first - 1));  // entry 0 landed
    wg_barrier();
    int g = 0;
    for (int t = 0; t < ntiles; ++t) {
      for (int kc = 0; kc < NKC; ++kc, ++g) {
        // the entry of interval g - 1 was read before the last barrier
        if (g >= 1 && g + kRing - 1 < nint) issue(g + kRing - 1);
        if (g + 1 < nint) {  // entry g + 1 lands before the barrier that opens it
          const int last = (g + kRing - 1 < nint) ? g + kRing - 1 : nint - 1;
          wait_vmcnt(kEntryLoads * (last - (g + 1)));
        }
        wg_barrier();
      }
    }
  } else {  // ---- phase wave: old phases of tile t into sP[t % 2]
    // the phases are read and written once per block: non-temporal loads
    // (cache policy 2) and stores keep them from displacing the waveform
    // chunks in L2 (K = 64: 8.8 -> 8.4 ms per block)
    int32_t sp16[16];  // spins of rows 8 j + l / 8 (padding slots read spin 0, never stored)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int64_t s = s0 + 8 * j + (lane >> 3);
      const int32_t v = slot_spin[s < nslots ? s : nslots - 1];
      sp16[j] = (s < nslots && v >= 0) ? v : 0;
    }
    const int pcol = ppiece_col(lane);
    auto issue = [&](int t) {
      float* dst = sP + (t & 1) * kTileWords;
      const int c = t * kTileW + pcol;
      const int cc = c < W ? c : t * kTileW;  // W % 4 == 0: a piece is all in or all out
#pragma unroll
      for (int j = 0; j < 16; ++j)
        __builtin_amdgcn_global_load_lds(phase + (int64_t)sp16[j] * W + cc, dst + 256 * j, 16, 0, 2);
    };
    issue(0);
    if (ntiles > 1) issue(1);
    wg_barrier();
    for (int t = 0; t < ntiles; ++t) {
      for (int kc = 0; kc < NKC; ++kc) {
        // sP[(t + 1) % 2] was last read in tile t - 1's epilogue, in the
        // first two chunk intervals of tile t
        if (kc == 2 && t >= 1 && t + 1 < ntiles) issue(t + 1);
        if (kc == NKC - 1) wait_vmcnt(t + 1 < ntiles ? 16 : 0);  // tile t landed
        wg_barrier();
      }
    }
  }
}

constexpr int kF32Chunk = 32;  // k per waveform chunk of the FP32 kernel

// The same contraction on the vector ALUs in FP32 (the CUDA-core form).
// Thread: one slot, 16 waveforms; the block's waveform columns via LDS.
__global__ void __launch_bounds__(256) phase_gemm_fp32_kernel(
    const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin, int32_t K3,
    const float* bt, int32_t W, float* phase) {
  __shared__ float sB[16][kF32Chunk];
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int w0 = blockIdx.y * 16;
  float acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.0f;
  for (int k0 = 0; k0 < K3; k0 += kF32Chunk) {
    for (int e = threadIdx.x; e < 16 * kF32Chunk; e += blockDim.x) {
      const int ww = e / kF32Chunk, kk = e % kF32Chunk;
      const int w = w0 + ww;
      sB[ww][kk] = (w < W && k0 + kk < K3) ? bt[(int64_t)w * K3 + k0 + kk] : 0.0f;
    }
    __syncthreads();
    if (s < nslots) {
      for (int kk = 0; kk < kF32Chunk && k0 + kk < K3; ++kk) {
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
    if (K3 != 96 && K3 != 192 && K3 != 288 && K3 != 384)
      return fail(-1, "nz_phase_gemm: MFMA mode needs K = 32, 64, 96 or 128");
    if (W % 4 || ld % 4) return fail(-1, "nz_phase_gemm: MFMA mode needs W and ld multiples of 4");
    dim3 grid((unsigned)((nslots + kTileS - 1) / kTileS));
    // K = 32, 64: one part; 96: three parts of 32 steps; 128: two parts of 64
    if (K3 == 96)
      hipLaunchKernelGGL((phase_gemm_mfma_kernel<3, 1, 32>), grid, dim3(64 * 6), 0, (hipStream_t)stream,
                         rec, ld, nslots, slot_spin, bt_hi, bt_lo, W, phase);
    else if (K3 == 192)
      hipLaunchKernelGGL((phase_gemm_mfma_kernel<3, 1, 64>), grid, dim3(64 * 6), 0, (hipStream_t)stream,
                         rec, ld, nslots, slot_spin, bt_hi, bt_lo, W, phase);
    else if (K3 == 288)
      hipLaunchKernelGGL((phase_gemm_mfma_kernel<3, 3, 32>), grid, dim3(64 * 14), 0, (hipStream_t)stream,
                         rec, ld, nslots, slot_spin, bt_hi, bt_lo, W, phase);
    else
      hipLaunchKernelGGL((phase_gemm_mfma_kernel<3, 2, 64>), grid, dim3(64 * 10), 0, (hipStream_t)stream,
                         rec, ld, nslots, slot_spin, bt_hi, bt_lo, W, phase);
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
