// Rejection seeding of spins inside target structures.
//
// Follows seedSpinsKernel (sim3d_kernel.cu:240-337) as driven by
// SimGeometry3D.gpu_seed (geometry.py:73-174): a uniform point of the box is
// accepted when it lies inside a target structure and inside no earlier
// non-target structure (structures are visited in id order, the box last).
// Equivalently: accepted iff the smallest id of a containing target structure
// is below the smallest id of a containing non-target structure.
//
// Differences by design: the reference relaunches the kernel with a fresh
// unseeded host random seed until every spin is placed and scans every sphere
// of every structure per attempt; here each spin loops over its own attempts
// in one launch (attempt k draws Philox counter (k, 0, spin, 1)), and the
// containment test reads only the point's cell list, which holds every sphere
// that can contain a point of that cell.
#include <hip/hip_runtime.h>

#include "nz_internal.h"
#include "nz_rng.h"

namespace nz {

constexpr int kMaxTargetWords = 256;  // up to 8192 structures in the kernel arguments

struct SeedArgs {
  Grid g;
  int64_t n;
  uint64_t seed;
  int32_t max_attempts;
  float* pos;
  int32_t* nfailed;
  uint32_t target_bits[kMaxTargetWords];
};

__device__ __forceinline__ bool is_target(const SeedArgs& A, int sid) {
  return (A.target_bits[sid >> 5] >> (sid & 31)) & 1u;
}

__global__ void __launch_bounds__(256) seed_kernel(const SeedArgs A) {
  const Grid& g = A.g;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const bool box_target = is_target(A, g.nstruct);
  for (int k = 0; k < A.max_attempts; ++k) {
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)k, 0u, (uint32_t)i, kStreamSeed},
                                  (uint32_t)A.seed, (uint32_t)(A.seed >> 32));
    float x[3];
    x[0] = (unif23(r.x) - 0.5f) * g.L[0];
    x[1] = (unif23(r.y) - 0.5f) * g.L[1];
    x[2] = (unif23(r.z) - 0.5f) * g.L[2];
    int l[3];
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double hd = (double)g.L[a] / g.ncell[a];
      int c = (int)floor(((double)x[a] + 0.5 * (double)g.L[a]) / hd);
      c = c < 0 ? 0 : (c > g.ncell[a] - 1 ? g.ncell[a] - 1 : c);
      l[a] = c + 1;
      p[a] = (float)((double)x[a] - (-0.5 * (double)g.L[a] + (double)c * hd));
    }
    // the point is strictly inside the box (geometry.py:69-71)
    int min_t = box_target ? g.nstruct : INT32_MAX;
    int min_n = box_target ? INT32_MAX : g.nstruct;
    const int cid = (l[2] * g.nloc[1] + l[1]) * g.nloc[0] + l[0];
    for (int q = g.cell_start[cid]; q < g.cell_start[cid + 1]; ++q) {
      const float4 S = g.cand[2 * (int64_t)q];
      const float4 B = g.cand[2 * (int64_t)q + 1];
      const float dx = p[0] - S.x, dy = p[1] - S.y, dz = p[2] - S.z;
      if (dx * dx + dy * dy + dz * dz < B.x) {  // sim3d_kernel.cu:299
        const int sid = (int)B.z;
        if (is_target(A, sid)) min_t = min(min_t, sid);
        else min_n = min(min_n, sid);
      }
    }
    if (min_t < min_n) {
      A.pos[i] = x[0];
      A.pos[A.n + i] = x[1];
      A.pos[2 * A.n + i] = x[2];
      return;
    }
  }
  if (A.nfailed) atomicAdd(A.nfailed, 1);
}

}  // namespace nz

using namespace nz;

extern "C" int nz_seed(nz_substrate_t sub, int64_t n, const int32_t* targets, int32_t ntargets,
                       uint64_t seed, int32_t max_attempts, float* pos, int32_t* nfailed,
                       void* stream) {
  if (!sub) return fail(-1, "nz_seed: substrate is NULL");
  if (n <= 0 || n > (int64_t)INT32_MAX) return fail(-1, "nz_seed: spin count must be in [1, 2^31)");
  if (!pos) return fail(-1, "nz_seed: pos is NULL");
  if (max_attempts <= 0) return fail(-1, "nz_seed: max_attempts must be > 0");
  const Substrate* s = (const Substrate*)sub;
  if (s->g.nstruct + 1 > kMaxTargetWords * 32)
    return fail(-1, "nz_seed: more than 8191 structures are not supported");
  SeedArgs args{};
  args.g = s->g;
  args.n = n;
  args.seed = seed;
  args.max_attempts = max_attempts;
  args.pos = pos;
  args.nfailed = nfailed;
  if (!targets || ntargets <= 0) {  // geometry.py:88-89: every structure and the box
    for (int k = 0; k <= s->g.nstruct; ++k) args.target_bits[k >> 5] |= 1u << (k & 31);
  } else {
    for (int q = 0; q < ntargets; ++q) {
      const int t = targets[q];
      if (t < 0 || t > s->g.nstruct) return fail(-1, "nz_seed: target structure id out of range");
      args.target_bits[t >> 5] |= 1u << (t & 31);
    }
  }
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(seed_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, args);
  NZ_HIP(hipGetLastError());
  return 0;
}
