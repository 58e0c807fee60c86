// The walker handle and its kernels: resident spin state, spatial sort, the
// persistent walk kernel (one spin per lane, many steps per launch; the
// per-spin logic is in nz_walk_core.h) and the read-back kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "nz_walk_core.h"

namespace nz {

// Per-slot resident state.  Slots are ordered by cell (see sort_spins);
// spin_id maps a slot to its spin, which keys the Philox stream and orders
// the read-back.
struct SlotState {
  int32_t* spin_id;  // [n]
  int32_t* ucell;    // [3,n] unwrapped cell index
  double* rel;       // [3,n] position relative to the cell's lower corner
  int32_t* ucell0;   // [3,n] start cell
  double* rel0;      // [3,n] start position relative to that cell
  int32_t* sid;      // [n]
  float* logsig;     // [n]
  float* mom;        // [3,n]
  int32_t* nsub;     // [n] or null
  int32_t* nguard;   // [n] or null
  uint32_t* tstep;   // [n] global index of the spin's next step
};

struct Walker {
  const Substrate* sub;
  int64_t n;
  uint32_t flags;
  SlotState buf[2];
  int cur = 0;
  int32_t* keys = nullptr;    // [n] sort keys
  int32_t* counts = nullptr;  // [nkeys + 1]
  int32_t* cursor = nullptr;  // [nkeys]
  int32_t* partial = nullptr; // [scan tiles] tile sums of the key-count scan
  float* pos_tmp = nullptr;   // [3,n] seeding output
  int64_t nkeys = 0;          // sort key range: 2 x cells (intra / extra)
  int32_t sort_interval = 64;
  int32_t xcd_chunk = 64;     // NZ_WALK_XCD_CHUNK in the environment overrides it
  int32_t since_sort = 1 << 30;
  int64_t step = 0;
  std::vector<void*> allocations;
  int device = 0;
  // walk_kernel launch timing (nz_walker_time_kernels): one start/stop event
  // pair per launch since the last nz_walker_kernel_time
  bool timing = false;
  std::vector<hipEvent_t> events;
  size_t events_used = 0;
};

struct WalkArgs {
  Grid g;
  SlotState st;
  int64_t n;
  uint64_t seed;
  int64_t base;      // global step of g[0] / garea[0] (the nz_walk call's first step)
  int64_t target;    // the wavefront runs until all its spins reached this step
  int64_t limit;     // a spin may run ahead up to this step
  int64_t mom_base;  // global step of stepmom[0] for this launch
  double dt;
  const float* gph;
  const float* garea;
  double* stepmom;
  uint32_t flags;
  int32_t xcd_chunk;  // slot blocks per XCD chunk (0: no remap)
};

// Workgroups of a launch are dispatched to the 8 XCDs round-robin (block b
// runs on XCD b % 8) and every XCD has its own L2.  The remap hands XCD x
// chunks of `chunk` consecutive (cell-sorted) slot blocks: within each group
// of 8*chunk blocks, XCD x processes blocks [x*chunk, (x+1)*chunk), so the
// blocks one L2 serves at a time cover a few compact regions of space
// instead of the whole span the chip is working on, while every XCD still
// gets an equal share of each part of the sorted order (the cost per spin
// differs between regions).  A final partial group keeps the identity map.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb, int64_t chunk) {
  constexpr int64_t kXcd = 8;
  if (chunk <= 0) return b;
  const int64_t group = kXcd * chunk;
  const int64_t g0 = (b / group) * group;
  if (g0 + group > nb) return b;
  return g0 + (b % kXcd) * chunk + (b - g0) / kXcd;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Steps per launch when per-step moment sums are requested (LDS partials).
constexpr int kMomChunk = 256;

// Each lane runs its own spin.  One iteration of the loop advances every
// lane by one sub-step; a lane whose step completes starts its next step in
// the following iteration, so lanes that scatter do not hold up the rest of
// the wavefront.  The wavefront runs until all its spins reached `target`;
// meanwhile spins that are ahead keep stepping up to `limit` (the end of the
// walk), so collision-poor spins do not idle at every launch boundary.
// Register budget of the walk kernel: waves per SIMD the compiler must fit
// (0 = unconstrained).  Three waves (168 VGPRs, a few spills) beat the
// unconstrained two (172 VGPRs) by 1.23x on the 200-axon substrate; four
// waves (128 VGPRs, more spills) measure the same as three.
#ifndef NZ_WALK_WAVES_PER_EU
#define NZ_WALK_WAVES_PER_EU 3
#endif
#if NZ_WALK_WAVES_PER_EU > 0
#define NZ_WALK_ATTR __attribute__((amdgpu_waves_per_eu(NZ_WALK_WAVES_PER_EU)))
#else
#define NZ_WALK_ATTR
#endif

__global__ void __launch_bounds__(256) NZ_WALK_ATTR walk_kernel(const WalkArgs A) {
  __shared__ float smom[kMomChunk * 6];
  const Grid& g = A.g;
  const SlotState& st = A.st;
  const int64_t n = A.n;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, A.xcd_chunk);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  const bool active = i < n;
  const int64_t ii = active ? i : 0;

  Lane s;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.u[a] = active ? st.ucell[a * n + ii] : 0;
    s.p[a] = active ? st.rel[a * n + ii] : 0.0;
    s.m[a] = (active && A.gph) ? st.mom[a * n + ii] : 0.0f;
  }
  lane_load_cells(g, s);
  s.logsig = active ? st.logsig[ii] : 0.0f;
  s.sid = active ? (float)st.sid[ii] : (float)g.nstruct;
  s.nsub = 0;
  s.nguard = 0;
  const bool do_t2 = (A.flags & NZ_WALK_T2) != 0;
  const uint32_t spin = active ? (uint32_t)st.spin_id[ii] : 0u;
  int64_t t = active ? (int64_t)st.tstep[ii] : A.target;
  int u0[3] = {0, 0, 0};
  double p0[3] = {0.0, 0.0, 0.0};
  if (A.stepmom) {
    for (int k = threadIdx.x; k < (int)(A.target - A.mom_base) * 6; k += blockDim.x) smom[k] = 0.0f;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      u0[a] = active ? st.ucell0[a * n + ii] : 0;
      p0[a] = active ? st.rel0[a * n + ii] : 0.0;
    }
  }

  bool started = false;
  StepState q;
  q.pending = -1;
  // A spin only stops at a step boundary: the sub-step state of a step in
  // progress is not kept across launches.
  while (__any(t < A.target || started)) {
    const bool behind = __any(t < A.target);
    if (active && (started || (t < A.limit && (t < A.target || behind)))) {
      const uint32_t step = (uint32_t)t;
      if (!started) {
        step_begin(g, s, q, A.dt);
        started = true;
      }
      if (q.pending >= 0) step_draw(q, A.seed, spin, step);
      if (substep(g, s, q, A.seed, spin, step, A.dt, do_t2)) {
        const float G = A.gph ? A.gph[t - A.base] : 0.0f;
        const float Ga = A.gph ? A.garea[t - A.base] : 0.0f;
        step_end(g, s, G, Ga);
        if (A.stepmom) {  // moment sums (sim3d_kernel.cu:191-237)
          const int k = (int)(t - A.mom_base) * 6;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const float d = (float)lane_disp(g, s, u0, p0, a);
            atomicAdd(&smom[k + a], d * d);
            atomicAdd(&smom[k + 3 + a], d * d * d * d);
          }
        }
        ++t;
        started = false;
      }
    }
  }
  if (A.stepmom) {  // block partials -> FP64 sums
    __syncthreads();
    for (int k = threadIdx.x; k < (int)(A.target - A.mom_base) * 6; k += blockDim.x)
      atomicAdd(&A.stepmom[(A.mom_base - A.base) * 6 + k], (double)smom[k]);
  }

  if (!active) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.ucell[a * n + i] = s.u[a];
    st.rel[a * n + i] = s.p[a];
    if (A.gph) st.mom[a * n + i] = s.m[a];
  }
  st.logsig[i] = s.logsig;
  st.sid[i] = (int32_t)s.sid;
  st.tstep[i] = (uint32_t)t;
  if (st.nsub) st.nsub[i] += s.nsub;
  if (st.nguard) st.nguard[i] += s.nguard;
}

// Absolute start positions (spin order) -> resident state (slot = spin).
__global__ void __launch_bounds__(256) init_kernel(const Grid g, const SlotState st, int64_t n,
                                                   const float* pos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Lane s;
  const float p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
  lane_init(g, p, s);
  st.spin_id[i] = (int32_t)i;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.ucell[a * n + i] = s.u[a];
    st.ucell0[a * n + i] = s.u[a];
    st.rel[a * n + i] = s.p[a];
    st.rel0[a * n + i] = s.p[a];
    st.mom[a * n + i] = 0.0f;
  }
  st.sid[i] = (int32_t)s.sid;
  st.logsig[i] = 0.0f;
  st.tstep[i] = 0u;
  if (st.nsub) st.nsub[i] = 0;
  if (st.nguard) st.nguard[i] = 0;
}

__global__ void fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- spatial counting sort by wrapped cell -------------------------------
__global__ void fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ int32_t wrapped_cell(const Grid& g, const int32_t* ucell, int64_t n,
                                                int64_t i) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int w = ucell[a * n + i] % g.ncell[a];
    c[a] = w < 0 ? w + g.ncell[a] : w;
  }
  return (c[2] * g.ncell[1] + c[1]) * g.ncell[0] + c[0];
}

// Runs of equal keys within a wavefront (the common case once spins are
// sorted) are counted with one atomic per run: lane `head` of the run holds
// its length.
struct WaveRun {
  int head;  // lane of the first lane of this lane's run
  int len;   // run length (meaningful on the head lane)
};

__device__ __forceinline__ WaveRun wave_run(int32_t key, bool valid) {
  const int lane = threadIdx.x & 63;
  const int32_t prev = __shfl_up(key, 1, 64);
  const bool start = valid && (lane == 0 || key != prev);
  const unsigned long long heads = __ballot(start);
  const unsigned long long vmask = __ballot(valid);
  // head: highest run start at or below this lane
  const unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
  WaveRun r;
  r.head = below ? 63 - __clzll(below) : lane;
  // run end: next run start above the head, or the end of the valid lanes
  const unsigned long long above = heads & ~((2ull << r.head) - 1 | (r.head == 63 ? ~0ull : 0ull));
  const int next = above ? __ffsll(above) - 1 : 64;
  const int last_valid = vmask ? 63 - __clzll(vmask) : -1;
  r.len = min(next, last_valid + 1) - r.head;
  return r;
}

__global__ void __launch_bounds__(256) sort_keys_kernel(const Grid g, const SlotState st, int64_t n,
                                                        int32_t* keys, int32_t* counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  // intra-structure spins first, then extra-structure spins; cell order
  // within each class (spins with similar collision rates share wavefronts)
  const int64_t ncells = (int64_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
  const int32_t k = valid ? wrapped_cell(g, st.ucell, n, i) +
                                (st.sid[i] >= g.nstruct ? (int32_t)ncells : 0)
                          : -1;
  if (valid) keys[i] = k;
  const WaveRun r = wave_run(k, valid);
  if (valid && r.head == (int)(threadIdx.x & 63)) atomicAdd(&counts[k], r.len);
}

// Exclusive scan of the m key counts in place, three coalesced passes over
// tiles of kScanTile: tile sums, a scan of the tile sums, and the scan of each
// tile from its offset.  The scan runs over m + 1 entries whose last one is 0,
// so counts[m] ends up holding the total.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanThreads * kScanPer;

// Exclusive scan of one value per thread across a 256-thread block; *total
// receives the block sum.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* total) {
  __shared__ int32_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    before += (w < wave) ? wsum[w] : 0;
    all += wsum[w];
  }
  __syncthreads();  // wsum may be reused by the caller's next scan
  *total = all;
  return before + x - v;
}

__global__ void __launch_bounds__(kScanThreads) tile_sum_kernel(const int32_t* counts, int64_t m,
                                                                int32_t* partial) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    s += (i < m) ? counts[i] : 0;
  }
  int32_t total;
  (void)block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive scan of the nb tile sums, one block, in chunks of 256.
__global__ void __launch_bounds__(kScanThreads) partial_scan_kernel(int32_t* partial, int64_t nb) {
  int32_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    const int64_t i = b0 + threadIdx.x;
    const int32_t v = (i < nb) ? partial[i] : 0;
    int32_t total;
    const int32_t ex = block_exclusive_scan(v, &total);
    if (i < nb) partial[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(kScanThreads) tile_scan_kernel(int32_t* counts, int64_t m,
                                                                 const int32_t* partial) {
  __shared__ int32_t tile[kScanTile + kScanTile / 32];  // +1 word per 32: fewer bank conflicts
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  auto slot = [](int k) { return k + (k >> 5); };
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int e = k * kScanThreads + threadIdx.x;
    const int64_t i = base + e;
    tile[slot(e)] = (i < m) ? counts[i] : 0;
  }
  __syncthreads();
  int32_t v[kScanPer], s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = tile[slot(threadIdx.x * kScanPer + k)];
    s += v[k];
  }
  int32_t total;
  int32_t run = block_exclusive_scan(s, &total) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    tile[slot(threadIdx.x * kScanPer + k)] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int e = k * kScanThreads + threadIdx.x;
    const int64_t i = base + e;
    if (i < m) counts[i] = tile[slot(e)];
  }
}

__global__ void __launch_bounds__(256) sort_scatter_kernel(const SlotState src, const SlotState dst,
                                                           int64_t n, const int32_t* keys,
                                                           const int32_t* offsets, int32_t* cursor) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const int32_t k = valid ? keys[i] : -1;
  const WaveRun r = wave_run(k, valid);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (valid && r.head == lane) base = atomicAdd(&cursor[k], r.len);
  base = __shfl(base, r.head, 64);
  if (!valid) return;
  const int64_t j = (int64_t)offsets[k] + base + (lane - r.head);
  dst.spin_id[j] = src.spin_id[i];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dst.ucell[a * n + j] = src.ucell[a * n + i];
    dst.rel[a * n + j] = src.rel[a * n + i];
    dst.ucell0[a * n + j] = src.ucell0[a * n + i];
    dst.rel0[a * n + j] = src.rel0[a * n + i];
    dst.mom[a * n + j] = src.mom[a * n + i];
  }
  dst.sid[j] = src.sid[i];
  dst.logsig[j] = src.logsig[i];
  dst.tstep[j] = src.tstep[i];
  if (src.nsub) dst.nsub[j] = src.nsub[i];
  if (src.nguard) dst.nguard[j] = src.nguard[i];
}

// ---- read-back in spin order ---------------------------------------------
__global__ void __launch_bounds__(256) read_kernel(const Grid g, const SlotState st, int64_t n,
                                                   float* pos, float* pos0, float* sig, float* mom,
                                                   double* pos_d, double* disp_d, int32_t* nsub,
                                                   int32_t* nguard) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = st.spin_id[i];
  if (k < 0 || k >= n) return;  // every slot holds a spin id in [0, n)
  Lane s;
  int u0[3];
  double p0[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.u[a] = st.ucell[a * n + i];
    s.p[a] = st.rel[a * n + i];
    u0[a] = st.ucell0[a * n + i];
    p0[a] = st.rel0[a * n + i];
  }
  lane_load_cells(g, s);
  if (pos || pos0) {
    float x[3], x0[3];
    lane_outputs(g, s, u0, p0, x, x0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (pos) pos[a * n + k] = x[a];
      if (pos0) pos0[a * n + k] = x0[a];
    }
  }
  if (sig) sig[k] = expf(st.logsig[i]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (mom) mom[a * n + k] = st.mom[a * n + i];
    if (pos_d) pos_d[a * n + k] = cell_lo(g, a, s.l[a]) + s.p[a];
    if (disp_d) disp_d[a * n + k] = lane_disp(g, s, u0, p0, a);
  }
  if (nsub) nsub[k] = st.nsub ? st.nsub[i] : 0;
  if (nguard) nguard[k] = st.nguard ? st.nguard[i] : 0;
}

__global__ void __launch_bounds__(256) disp_moments_kernel(const Grid g, const SlotState st,
                                                           int64_t n, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float d = (float)((double)(st.ucell[a * n + i] - st.ucell0[a * n + i]) * g.hd[a] +
                              (st.rel[a * n + i] - st.rel0[a * n + i]));
      v[a] = d * d;
      v[3 + a] = v[a] * v[a];
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float s = wave_sum(v[q]);
    if ((threadIdx.x & 63) == 0) atomicAdd(&out[q], (double)s);
  }
}

__global__ void zero_f64_kernel(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

static unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

static int walker_alloc(Walker* w, size_t bytes, void** out) {
  void* p = nullptr;
  NZ_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16)));
  w->allocations.push_back(p);
  *out = p;
  return 0;
}

static void walker_release(Walker* w) {
  for (void* p : w->allocations) (void)hipFree(p);
  w->allocations.clear();
  for (hipEvent_t e : w->events) (void)hipEventDestroy(e);
  w->events.clear();
  w->events_used = 0;
}

// The event pair of the next timed launch (created on first use).
static int next_event_pair(Walker* w, hipEvent_t* start, hipEvent_t* stop) {
  while (w->events.size() < w->events_used + 2) {
    hipEvent_t e;
    NZ_HIP(hipEventCreate(&e));
    w->events.push_back(e);
  }
  *start = w->events[w->events_used];
  *stop = w->events[w->events_used + 1];
  w->events_used += 2;
  return 0;
}

static int sort_spins(Walker* w, hipStream_t stream) {
  const Grid& g = w->sub->g;
  const int64_t n = w->n;
  hipLaunchKernelGGL(fill_i32, dim3(blocks_for(w->nkeys + 1)), dim3(256), 0, stream, w->counts,
                     w->nkeys + 1, 0);
  hipLaunchKernelGGL(fill_i32, dim3(blocks_for(w->nkeys)), dim3(256), 0, stream, w->cursor,
                     w->nkeys, 0);
  hipLaunchKernelGGL(sort_keys_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, g, w->buf[w->cur],
                     n, w->keys, w->counts);
  const int64_t m = w->nkeys + 1, nb = (m + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(tile_sum_kernel, dim3(nb), dim3(kScanThreads), 0, stream, w->counts, m,
                     w->partial);
  hipLaunchKernelGGL(partial_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, w->partial, nb);
  hipLaunchKernelGGL(tile_scan_kernel, dim3(nb), dim3(kScanThreads), 0, stream, w->counts, m,
                     w->partial);
  hipLaunchKernelGGL(sort_scatter_kernel, dim3(blocks_for(n)), dim3(256), 0, stream,
                     w->buf[w->cur], w->buf[1 - w->cur], n, w->keys, w->counts, w->cursor);
  NZ_HIP(hipGetLastError());
  w->cur = 1 - w->cur;
  w->since_sort = 0;
  return 0;
}

static int check_walker(const Walker* w, const char* who) {
  if (!w) return fail(-1, std::string(who) + ": walker is NULL");
  return 0;
}

}  // namespace nz

using namespace nz;

extern "C" {

int nz_walker_create(nz_substrate_t sub, int64_t n, uint32_t flags, nz_walker_t* out) {
  if (!out) return fail(-1, "nz_walker_create: out is NULL");
  *out = nullptr;
  if (!sub) return fail(-1, "nz_walker_create: substrate is NULL");
  if (n <= 0 || n > (int64_t)INT32_MAX) return fail(-1, "nz_walker_create: spin count must be in [1, 2^31)");
  Walker* w = new Walker();
  w->sub = (const Substrate*)sub;
  w->n = n;
  w->flags = flags;
  (void)hipGetDevice(&w->device);
  const Grid& g = w->sub->g;
  w->nkeys = 2 * (int64_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
  if (const char* e = getenv("NZ_WALK_XCD_CHUNK")) w->xcd_chunk = std::max(0, atoi(e));
  int rc = 0;
  for (int b = 0; b < 2 && !rc; ++b) {
    SlotState& s = w->buf[b];
    void* p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * n, &p))) break;
    s.spin_id = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * 3 * n, &p))) break;
    s.ucell = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(double) * 3 * n, &p))) break;
    s.rel = (double*)p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * 3 * n, &p))) break;
    s.ucell0 = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(double) * 3 * n, &p))) break;
    s.rel0 = (double*)p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * n, &p))) break;
    s.sid = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(float) * n, &p))) break;
    s.logsig = (float*)p;
    if ((rc = walker_alloc(w, sizeof(float) * 3 * n, &p))) break;
    s.mom = (float*)p;
    if ((rc = walker_alloc(w, sizeof(uint32_t) * n, &p))) break;
    s.tstep = (uint32_t*)p;
    s.nsub = s.nguard = nullptr;
    if (flags & NZ_WALKER_COUNTERS) {
      if ((rc = walker_alloc(w, sizeof(int32_t) * n, &p))) break;
      s.nsub = (int32_t*)p;
      if ((rc = walker_alloc(w, sizeof(int32_t) * n, &p))) break;
      s.nguard = (int32_t*)p;
    }
  }
  void* p = nullptr;
  if (!rc) rc = walker_alloc(w, sizeof(int32_t) * n, &p);
  if (!rc) { w->keys = (int32_t*)p; rc = walker_alloc(w, sizeof(int32_t) * (w->nkeys + 1), &p); }
  if (!rc) { w->counts = (int32_t*)p; rc = walker_alloc(w, sizeof(int32_t) * w->nkeys, &p); }
  if (!rc) {
    w->cursor = (int32_t*)p;
    rc = walker_alloc(w, sizeof(int32_t) * ((w->nkeys + 1 + kScanTile - 1) / kScanTile), &p);
  }
  if (!rc) { w->partial = (int32_t*)p; rc = walker_alloc(w, sizeof(float) * 3 * n, &p); }
  if (!rc) w->pos_tmp = (float*)p;
  if (rc) {
    walker_release(w);
    delete w;
    return rc;
  }
  *out = (nz_walker_t)w;
  return 0;
}

int nz_walker_destroy(nz_walker_t ww) {
  if (!ww) return 0;
  Walker* w = (Walker*)ww;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(w->device);
  walker_release(w);
  (void)hipSetDevice(dev);
  delete w;
  return 0;
}

int nz_walker_set_sort_interval(nz_walker_t ww, int32_t steps) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_set_sort_interval")) return rc;
  if (steps < 0) return fail(-1, "nz_walker_set_sort_interval: steps < 0");
  w->sort_interval = steps;
  return 0;
}

int nz_walker_get_step(nz_walker_t ww, int64_t* step) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_get_step")) return rc;
  if (!step) return fail(-1, "nz_walker_get_step: step is NULL");
  *step = w->step;
  return 0;
}

int nz_walker_set_step(nz_walker_t ww, int64_t step) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_set_step")) return rc;
  if (step < 0 || step > (int64_t)UINT32_MAX) return fail(-1, "nz_walker_set_step: out of range");
  w->step = step;
  hipLaunchKernelGGL(fill_u32, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)0,
                     w->buf[w->cur].tstep, w->n, (uint32_t)step);
  NZ_HIP(hipGetLastError());
  NZ_HIP(hipStreamSynchronize((hipStream_t)0));
  return 0;
}

int nz_walker_set_positions(nz_walker_t ww, const float* pos, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_set_positions")) return rc;
  if (!pos) return fail(-1, "nz_walker_set_positions: pos is NULL");
  w->cur = 0;
  hipLaunchKernelGGL(init_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[0], w->n, pos);
  NZ_HIP(hipGetLastError());
  w->step = 0;
  w->since_sort = 1 << 30;
  return 0;
}

int nz_walker_seed(nz_walker_t ww, const int32_t* targets, int32_t ntargets, uint64_t seed,
                   int32_t max_attempts, int32_t* nfailed, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_seed")) return rc;
  if (int rc = nz_seed((nz_substrate_t)w->sub, w->n, targets, ntargets, seed, max_attempts,
                       w->pos_tmp, nfailed, stream))
    return rc;
  return nz_walker_set_positions(ww, w->pos_tmp, stream);
}

int nz_walk(nz_walker_t ww, uint64_t seed, int32_t nsteps, double dt, const float* g,
            const float* garea, double* stepmom, uint32_t flags, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walk")) return rc;
  if (nsteps < 0) return fail(-1, "nz_walk: nsteps < 0");
  if (!(dt >= 0.0)) return fail(-1, "nz_walk: dt must be >= 0");
  if ((g == nullptr) != (garea == nullptr)) return fail(-1, "nz_walk: pass both g and garea or neither");
  if (w->step + nsteps > (int64_t)UINT32_MAX) return fail(-1, "nz_walk: step counter overflow");
  hipStream_t s = (hipStream_t)stream;
  const int64_t base = w->step, end = w->step + nsteps;
  int64_t target = base;
  while (target < end) {
    if (w->sort_interval > 0 && w->since_sort >= w->sort_interval)
      if (int rc = sort_spins(w, s)) return rc;
    int64_t chunk = end - target;
    if (w->sort_interval > 0) chunk = std::min<int64_t>(chunk, w->sort_interval - w->since_sort);
    if (stepmom) chunk = std::min<int64_t>(chunk, kMomChunk);
    WalkArgs a;
    a.g = w->sub->g;
    a.st = w->buf[w->cur];
    a.n = w->n;
    a.seed = seed;
    a.base = base;
    a.mom_base = target;
    a.target = target + chunk;
    a.limit = stepmom ? a.target : end;  // per-step sums need the wavefront in step
    a.dt = dt;
    a.gph = g;
    a.garea = garea;
    a.stepmom = stepmom;
    a.flags = flags;
    a.xcd_chunk = w->xcd_chunk;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (w->timing) {
      if (int rc = next_event_pair(w, &e0, &e1)) return rc;
      NZ_HIP(hipEventRecord(e0, s));
    }
    hipLaunchKernelGGL(walk_kernel, dim3(blocks_for(w->n)), dim3(256), 0, s, a);
    NZ_HIP(hipGetLastError());
    if (w->timing) NZ_HIP(hipEventRecord(e1, s));
    target += chunk;
    w->since_sort += (int32_t)chunk;
  }
  w->step = end;
  return 0;
}

int nz_walker_read(nz_walker_t ww, float* pos, float* pos0, float* sig, float* mom, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_read")) return rc;
  hipLaunchKernelGGL(read_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[w->cur], w->n, pos, pos0, sig, mom, (double*)nullptr,
                     (double*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_walker_read_f64(nz_walker_t ww, double* pos, double* disp, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_read_f64")) return rc;
  hipLaunchKernelGGL(read_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[w->cur], w->n, (float*)nullptr, (float*)nullptr,
                     (float*)nullptr, (float*)nullptr, pos, disp, (int32_t*)nullptr,
                     (int32_t*)nullptr);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_walker_counters(nz_walker_t ww, int32_t* nsub, int32_t* nguard, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_counters")) return rc;
  if (!(w->flags & NZ_WALKER_COUNTERS))
    return fail(-1, "nz_walker_counters: walker created without NZ_WALKER_COUNTERS");
  hipLaunchKernelGGL(read_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[w->cur], w->n, (float*)nullptr, (float*)nullptr,
                     (float*)nullptr, (float*)nullptr, (double*)nullptr, (double*)nullptr, nsub,
                     nguard);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_walker_time_kernels(nz_walker_t ww, int32_t enable) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_time_kernels")) return rc;
  w->timing = enable != 0;
  w->events_used = 0;
  return 0;
}

int nz_walker_kernel_time(nz_walker_t ww, double* ms_total, int64_t* launches) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_kernel_time")) return rc;
  if (!ms_total || !launches) return fail(-1, "nz_walker_kernel_time: NULL output");
  double total = 0.0;
  for (size_t k = 0; k + 1 < w->events_used; k += 2) {
    NZ_HIP(hipEventSynchronize(w->events[k + 1]));
    float ms = 0.0f;
    NZ_HIP(hipEventElapsedTime(&ms, w->events[k], w->events[k + 1]));
    total += ms;
  }
  *ms_total = total;
  *launches = (int64_t)(w->events_used / 2);
  w->events_used = 0;
  return 0;
}

int nz_displacement_moments(nz_walker_t ww, double* out, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_displacement_moments")) return rc;
  if (!out) return fail(-1, "nz_displacement_moments: out is NULL");
  hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, out, (int64_t)6);
  hipLaunchKernelGGL(disp_moments_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[w->cur], w->n, out);
  NZ_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
