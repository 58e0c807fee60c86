// Phase projection and DWI signal reduction.
//
// The reference accumulates one phase per (direction, spin) at every step
// (randomWalk3d_phase_multidirr, sim3d_kernel.cu:880-942): 8*W bytes of
// read-modify-write per spin-step.  All of its directions share one gradient
// waveform G_t per launch, so phase[w] = u_w . sum_t G_t x_t (+ the same
// projection of the wrap corrections): the walker keeps that 3-vector moment
// and the W phases are formed once, here.  The signal step (sum over spins of
// exp(i*phase)) has no counterpart in the reference, whose DWI caller is
// missing; the sums are accumulated in FP64, per block and then over the
// blocks in a fixed order, so a given input always gives the same bits.
#include <hip/hip_runtime.h>

#include <mutex>

#include "nz_internal.h"

namespace nz {

__global__ void __launch_bounds__(256) project_kernel(int64_t n, int32_t W, const float* mom,
                                                      const float* dirs, const float* scale,
                                                      float* phase) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (i >= n) return;
  const float s = scale ? scale[w] : 1.0f;
  const float ux = dirs[w], uy = dirs[W + w], uz = dirs[2 * W + w];
  phase[(int64_t)w * n + i] = s * (ux * mom[i] + uy * mom[n + i] + uz * mom[2 * n + i]);
}

constexpr int kSigTile = 256;       // spins staged per LDS tile
constexpr int kSigMaxBlocks = 1024; // blocks, i.e. FP64 partial sums per signal

// Blocks of a reduction over n spins and the tiles each walks: the partition
// depends on n only, so a given n is always summed in the same order.
static void signal_partition(int64_t n, int64_t* nb, int64_t* tiles_per_block) {
  const int64_t ntiles = (n + kSigTile - 1) / kSigTile;
  *nb = ntiles < kSigMaxBlocks ? ntiles : kSigMaxBlocks;
  *tiles_per_block = (ntiles + *nb - 1) / *nb;
}

// Block b sums tiles [b tpb, (b + 1) tpb) of the spins.  Thread t handles
// signals t, t+256, ... reading every staged moment from LDS (broadcast
// reads).  mom is [3, ld]; entries whose valid[i] < 0 (padding slots of a
// batch) count as absent.  The block's FP64 sums go to partial[b][2W]
// (cos, sin interleaved), every entry written.
__global__ void __launch_bounds__(256) signal_kernel(int64_t n, int64_t ld, const int32_t* valid,
                                                     int32_t W, const float* mom,
                                                     const float* dirs, const float* scale,
                                                     int64_t tiles_per_block, double* partial) {
  __shared__ float sm[3][kSigTile];
  __shared__ int wcnt[kSigTile / 64];
  const int64_t base = (int64_t)blockIdx.x * tiles_per_block * kSigTile;
  constexpr int kMaxPer = 8;  // signals per thread (W <= 2048)
  double ac[kMaxPer], as[kMaxPer];
  float ux[kMaxPer], uy[kMaxPer], uz[kMaxPer];
#pragma unroll
  for (int q = 0; q < kMaxPer; ++q) {
    ac[q] = 0.0;
    as[q] = 0.0;
    const int w = threadIdx.x + q * 256;
    const float s = (w < W) ? (scale ? scale[w] : 1.0f) : 0.0f;
    ux[q] = (w < W) ? s * dirs[w] : 0.0f;
    uy[q] = (w < W) ? s * dirs[W + w] : 0.0f;
    uz[q] = (w < W) ? s * dirs[2 * W + w] : 0.0f;
  }
  for (int64_t tile = 0; tile < tiles_per_block; ++tile) {
    const int64_t t0 = base + tile * kSigTile;
    if (t0 >= n) break;
    const int64_t i = t0 + threadIdx.x;
    __syncthreads();
    const bool ok = i < n && (!valid || valid[i] >= 0);
    const float v0 = ok ? mom[i] : 0.0f;
    const float v1 = ok ? mom[ld + i] : 0.0f;
    const float v2 = ok ? mom[2 * ld + i] : 0.0f;
    // the present entries are packed to the front of the tile (ballot + count)
    const unsigned long long m = __ballot(ok);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int before = 0, cnt = 0;
    for (int q = 0; q < kSigTile / 64; ++q) {
      before += (q < wave) ? wcnt[q] : 0;
      cnt += wcnt[q];
    }
    if (ok) {
      const int dst = before + __popcll(m & ((1ull << lane) - 1));
      sm[0][dst] = v0;
      sm[1][dst] = v1;
      sm[2][dst] = v2;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxPer; ++q) {
      if ((int)threadIdx.x + q * 256 >= W) break;
      float pc = 0.0f, ps = 0.0f;  // partial sums of 32 spins in FP32, then FP64
      for (int k = 0; k < cnt; ++k) {
        const float ph = ux[q] * sm[0][k] + uy[q] * sm[1][k] + uz[q] * sm[2][k];
        float s, c;
        sincosf(ph, &s, &c);
        pc += c;
        ps += s;
        if ((k & 31) == 31) {
          ac[q] += (double)pc;
          as[q] += (double)ps;
          pc = 0.0f;
          ps = 0.0f;
        }
      }
      ac[q] += (double)pc;
      as[q] += (double)ps;
    }
  }
  double* out = partial + (int64_t)blockIdx.x * 2 * W;
#pragma unroll
  for (int q = 0; q < kMaxPer; ++q) {
    const int w = threadIdx.x + q * 256;
    if (w < W) {
      out[2 * w] = ac[q];
      out[2 * w + 1] = as[q];
    }
  }
}

// out[e] = sum over the nb blocks of partial[b][e], one workgroup per e:
// thread t adds blocks t, t + 256, ... in order, then a fixed tree over the
// 256 thread sums (the same order on every call).
__global__ void __launch_bounds__(256) signal_sum_kernel(const double* partial, int64_t nb,
                                                         int32_t W, double* out) {
  __shared__ double red[256];
  const int64_t e = blockIdx.x;
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) s += partial[b * 2 * W + e];
  red[threadIdx.x] = s;
  __syncthreads();
#pragma unroll
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e] = red[0];
}

int64_t signal_partial_len(int64_t n, int32_t W) {
  int64_t nb, tpb;
  signal_partition(n, &nb, &tpb);
  return nb * 2 * (int64_t)W;
}

int signal_reduce(int64_t n, int64_t ld, const int32_t* valid, int32_t W, const float* mom,
                  const float* dirs, const float* scale, double* partial, double* out,
                  hipStream_t stream) {
  if (n <= 0 || W <= 0 || W > 2048) return fail(-1, "signal reduction: need n > 0 and 0 < W <= 2048");
  if (!mom || !dirs || !out || !partial) return fail(-1, "signal reduction: NULL buffer");
  int64_t nb, tpb;
  signal_partition(n, &nb, &tpb);
  hipLaunchKernelGGL(signal_kernel, dim3((unsigned)nb), dim3(256), 0, stream, n, ld, valid, W, mom,
                     dirs, scale, tpb, partial);
  hipLaunchKernelGGL(signal_sum_kernel, dim3((unsigned)(2 * W)), dim3(256), 0, stream,
                     (const double*)partial, nb, W, out);
  NZ_HIP(hipGetLastError());
  return 0;
}

// Signal reduction with library-allocated partials, taken from the stream's
// memory pool and returned to it after the two kernels (stream-ordered: no
// synchronisation).
int signal_reduce_pooled(int64_t n, int64_t ld, const int32_t* valid, int32_t W, const float* mom,
                         const float* dirs, const float* scale, double* out, hipStream_t stream) {
  if (n <= 0 || W <= 0 || W > 2048) return fail(-1, "signal reduction: need n > 0 and 0 < W <= 2048");
  void* partial = nullptr;
  NZ_HIP(hipMallocAsync(&partial, sizeof(double) * (size_t)signal_partial_len(n, W), stream));
  const int rc = signal_reduce(n, ld, valid, W, mom, dirs, scale, (double*)partial, out, stream);
  const hipError_t e = hipFreeAsync(partial, stream);
  if (rc != 0) return rc;
  NZ_HIP(e);
  return 0;
}

// ---- allocation-free variant ------------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid it): the partials live
// in a file-scope device array, one copy per GPU, held until the process
// ends.  Its host-side name is not a device address: the address of each
// GPU's copy is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that pointer is handed to the kernels.
constexpr int32_t kNoallocMaxW = 512;
constexpr int64_t kNoallocPartials = (int64_t)kSigMaxBlocks * 2 * kNoallocMaxW;  // 8 MiB
static __device__ __align__(256) double scratch[kNoallocPartials];

constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

// This device's copy of `scratch` (looked up on the first call on a device).
static int scratch_of_current_device(double** out) {
  int dev = 0;
  NZ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(-1, "nz_signal_reduce_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    NZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(-1, "nz_signal_reduce_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace nz

using namespace nz;

extern "C" {

int nz_project_phase(int64_t n, int32_t W, const float* mom, const float* dirs, const float* scale,
                     float* phase, void* stream) {
  if (n <= 0 || W <= 0 || W > 65535) return fail(-1, "nz_project_phase: need n > 0 and 0 < W <= 65535");
  if (!mom || !dirs || !phase) return fail(-1, "nz_project_phase: NULL buffer");
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)W);
  hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, (hipStream_t)stream, n, W, mom, dirs, scale,
                     phase);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_signal_reduce(int64_t n, int32_t W, const float* mom, const float* dirs, const float* scale,
                     double* out, void* stream) {
  if (n <= 0 || W <= 0 || W > 2048) return fail(-1, "nz_signal_reduce: need n > 0 and 0 < W <= 2048");
  return signal_reduce_pooled(n, n, nullptr, W, mom, dirs, scale, out, (hipStream_t)stream);
}

int nz_signal_reduce_noalloc(int64_t n, int32_t W, const float* mom, const float* dirs,
                             const float* scale, double* out, void* stream) {
  if (n <= 0 || W <= 0 || W > kNoallocMaxW)
    return fail(-1, "nz_signal_reduce_noalloc: need n > 0 and 0 < W <= 512 (the static scratch "
                    "holds 1024 blocks x 2W FP64 partials)");
  if (!mom || !dirs || !out) return fail(-1, "nz_signal_reduce_noalloc: NULL buffer");
  if (signal_partial_len(n, W) > kNoallocPartials)
    return fail(-1, "nz_signal_reduce_noalloc: partials exceed the static scratch");
  double* partial = nullptr;
  if (int rc = scratch_of_current_device(&partial)) return rc;
  return signal_reduce(n, n, nullptr, W, mom, dirs, scale, partial, out, (hipStream_t)stream);
}

}  // extern "C"
