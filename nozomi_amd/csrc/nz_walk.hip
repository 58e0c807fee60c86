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
  int32_t* cell;     // [3,n] local index of the cell holding x
  double* x;         // [3,n] absolute position (wrapped as the reference wraps)
  int32_t* wrap;     // [3,n] net +L shifts of the start position
  double* x0;        // [3,n] start position
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
  int32_t sort_interval = 256;
  int32_t xcd_chunk = 64;     // NZ_WALK_XCD_CHUNK in the environment overrides it
  // threads per walk_kernel workgroup (NZ_WALK_BLOCK: 64, 128, 256).  One wave
  // per workgroup: a wave that finishes its spins frees its slot at once
  // instead of waiting for the three others of a 256-thread workgroup (200-axon
  // substrate, 1e6 spins, 250-step launches: 13.37 ms against 13.95 ms).
  int32_t walk_block = 64;
  // threads per moments_kernel workgroup (NZ_TRACK_BLOCK); 256 measured faster here than 64
  // (the per-step moment partials are per workgroup)
  int32_t track_block = 256;
  int32_t since_sort = 1 << 30;
  int64_t step = 0;
  // the last record call let spins run up to rec_ahead steps past its block
  // (nz_walk_record_ring); those steps sit in its ring, which the next record
  // block continues at entry ring_next
  int32_t rec_ahead = 0;
  int32_t ring_next = 0;
  int32_t ring = 0;
  int64_t rec_ld = 0;
  const float* rec_ptr = nullptr;
  std::vector<void*> allocations;
  int device = 0;
  // walk_kernel launch timing (nz_walker_time_kernels): one start/stop event
  // pair per launch since the last nz_walker_kernel_time
  bool timing = false;
  std::vector<hipEvent_t> events;
  size_t events_used = 0;
};

struct WalkArgs {
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
  float* rec;         // displacement record (nz_walk_record) or null
  int64_t rec_ld;     // floats per slot of the record
  int32_t rec_ring = 0;  // record entries per slot, used as a ring of steps
  int32_t rec_pos0 = 0;  // ring entry of the launch's first step (mom_base)
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

__device__ __forceinline__ double wave_sum_d(double v) {
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
// (0 = unconstrained).  On the 200-axon substrate (250-step launches over
// 1e6 spins) four waves (128 VGPRs, 28 bytes per lane of spills) run 13.9 ms
// per launch, three (140 VGPRs, none) 15.3 ms, five (96 VGPRs, 176 bytes of
// spills) 20.7 ms.
#ifndef NZ_WALK_WAVES_PER_EU
#define NZ_WALK_WAVES_PER_EU 4
#endif
#if NZ_WALK_WAVES_PER_EU > 0
#define NZ_WALK_ATTR __attribute__((amdgpu_waves_per_eu(NZ_WALK_WAVES_PER_EU)))
#else
#define NZ_WALK_ATTR
#endif
// The tracking walk (displacement record of the precision form, per-step
// moments of the narrow-pulse driver) also holds the start position.
#ifndef NZ_TRACK_WAVES_PER_EU
#define NZ_TRACK_WAVES_PER_EU 4
#endif
#if NZ_TRACK_WAVES_PER_EU > 0
#define NZ_TRACK_ATTR __attribute__((amdgpu_waves_per_eu(NZ_TRACK_WAVES_PER_EU)))
#else
#define NZ_TRACK_ATTR
#endif

// Per-launch parameters of one simulation's walk.
// Step numbers are 32-bit on the device (the host keeps them below 2^32).
struct StepParams {
  uint64_t seed;
  uint32_t base;      // global step of gph[0] / garea[0]
  uint32_t target;    // the wavefront runs until all its spins reached this step
  uint32_t limit;     // a spin may run ahead up to this step
  uint32_t mom_base;  // global step of stepmom[0] for this launch
  double dt;
  const float* gph;
  const float* garea;
  double* stepmom;   // per-step moment sums (needs smom and target - mom_base <= kMomChunk)
  uint32_t flags;
  float* rec;        // per-step displacement record [slots][rec_ld] or null
  int64_t rec_ld;    // floats per slot of the record (>= 3 * rec_ring)
  uint32_t rec_ring = 0;  // step t goes to ring entry (t - mom_base + rec_pos0) mod rec_ring
  uint32_t rec_pos0 = 0;
};

// What a walk launch keeps beside the walk itself.  The tracking kinds hold
// the start position in registers; each kind is its own kernel, so the plain
// walk and the record leave the other kinds' registers to the scan.
enum WalkKind { kPlain = 0, kMoments = 1, kRecord = 2 };

// LDS words read volatile: at each use, not kept in registers.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const volatile f64x2 lds_f64x2;

// Walk the spin held in slot i (fields strided by ld) from its next step to
// at least P.target.  Every lane of the block calls it (inactive lanes too:
// the per-step moment partials synchronise the block).  kMoments: per-step
// FP64 moment sums (P.stepmom); kRecord: per-step displacement record (P.rec).
// xs: the tracking kinds' LDS words, 4 doubles per thread.
template <int kKind>
__device__ __forceinline__ void walk_slot(const Grid& g, const SlotState& st, int64_t ld,
                                          int64_t i, bool active, const StepParams P,
                                          double* smom, double* xs = nullptr) {
  const ListBufs lb = list_bufs(g);  // every lane is active here
  // the recording walk keeps no phase moment and no T2 weight (the precision
  // form takes neither), leaving their registers to the walk
  constexpr bool kTrack = kKind != kRecord;
  const int64_t ii = active ? i : 0;
  Lane s;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.x[a] = active ? st.x[a * ld + ii] : 0.0;
    s.l[a] = active ? st.cell[a * ld + ii] : 1;
    s.w[a] = active ? st.wrap[a * ld + ii] : 0;
    s.m[a] = (kTrack && active && P.gph) ? st.mom[a * ld + ii] : 0.0f;
  }
  s.logsig = (kTrack && active) ? st.logsig[ii] : 0.0f;
  s.sid = active ? (float)st.sid[ii] : (float)g.nstruct;
  s.nsub = 0;
  s.nguard = 0;
  const bool do_t2 = kTrack && (P.flags & NZ_WALK_T2) != 0;
  const uint32_t spin = active ? (uint32_t)st.spin_id[ii] : 0u;
  uint32_t t = active ? st.tstep[ii] : P.target;
  if (kKind == kMoments) {
    for (int k = threadIdx.x; k < (int)(P.target - P.mom_base) * 6; k += blockDim.x) smom[k] = 0.0;
    __syncthreads();
  }
  // The tracking kinds keep the start position in LDS (xs: 4 doubles per
  // thread; kRecord adds the record row's offset as the fourth) and read it
  // at each step end; held in registers it was spilled to scratch, and each
  // reload waited behind the wave's outstanding vector loads
  if (kKind != kPlain) {
#pragma unroll
    for (int a = 0; a < 3; ++a) xs[4 * threadIdx.x + a] = active ? st.x0[a * ld + ii] : 0.0;
    xs[4 * threadIdx.x + 3] =
        __longlong_as_double(kKind == kRecord ? (long long)(ii * P.rec_ld) : 0ll);
  }

  bool started = false;
  StepState q;
  q.pending = -1;
  q.ds_sid = -1.0f;
  // A spin only stops at a step boundary: the sub-step state of a step in
  // progress is not kept across launches.  (The wavefront's exit test sits
  // at the bottom of the loop: tested at the top, the compiler copies the
  // lane state in and out of the loop every iteration.)
  bool go = __any(t < P.target);
  while (go) {
    const bool behind = __any(t < P.target);
    if (active && (started || (t < P.limit && (t < P.target || behind)))) {
      const uint32_t step = t;
      if (!started) {
        step_begin(g, s, q, P.dt);
        started = true;
      }
      if (q.pending >= 0) step_draw(q, P.seed, spin, step);
      // the step's gradient samples, loaded before the sub-step so that their
      // latency hides behind it (used only if the sub-step ends the step)
      float G = 0.0f, Ga = 0.0f;
      if (kTrack && P.gph) {
        G = P.gph[t - P.base];
        Ga = P.garea[t - P.base];
      }
      sched_fence();
      if (substep(g, lb, s, q, P.seed, spin, step, P.dt, do_t2)) {
        step_end(g, s, G, Ga, kTrack);
        if (kKind == kRecord) {  // displacement since the start, for the blocked phase contraction
          // slot-major: a lane's steps are contiguous (the lanes of a
          // wavefront are at different steps, so rows of steps would scatter)
          uint32_t e = t - P.mom_base + P.rec_pos0;  // < 2 rec_ring
          if (e >= P.rec_ring) e -= P.rec_ring;
          lds_f64x2* xv = (lds_f64x2*)(xs + 4 * threadIdx.x);
          const f64x2 u0 = xv[0], u1 = xv[1];
          const double xl[3] = {u0.x, u0.y, u1.x};
          float* r = P.rec + __double_as_longlong(u1.y) + (int64_t)e * 3;
          r[0] = (float)lane_disp(g, s, xl, 0);
          r[1] = (float)lane_disp(g, s, xl, 1);
          r[2] = (float)lane_disp(g, s, xl, 2);
        }
        if (kKind == kMoments) {  // moment sums (sim3d_kernel.cu:191-237), FP64 throughout
          const int k = (int)(t - P.mom_base) * 6;
          const f64x2 u0 = ((lds_f64x2*)(xs + 4 * threadIdx.x))[0];
          const double x0z = ((lds_f64x2*)(xs + 4 * threadIdx.x))[1].x;
          const double xl[3] = {u0.x, u0.y, x0z};
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const double d = lane_disp(g, s, xl, a);
            const double d2 = d * d;
            atomicAdd(&smom[k + a], d2);
            atomicAdd(&smom[k + 3 + a], d2 * d2);
          }
        }
        ++t;
        started = false;
      }
    }
    go = __any(t < P.target || started);
  }
  if (kKind == kMoments) {  // block partials -> FP64 sums
    __syncthreads();
    for (int k = threadIdx.x; k < (int)(P.target - P.mom_base) * 6; k += blockDim.x)
      atomicAdd(&P.stepmom[(int64_t)(P.mom_base - P.base) * 6 + k], smom[k]);
  }

  if (!active) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.x[a * ld + i] = s.x[a];
    st.cell[a * ld + i] = s.l[a];
    st.wrap[a * ld + i] = s.w[a];
    if (kTrack && P.gph) st.mom[a * ld + i] = s.m[a];
  }
  if (kTrack) st.logsig[i] = s.logsig;
  st.sid[i] = (int32_t)s.sid;
  st.tstep[i] = t;
  if (st.nsub) st.nsub[i] += s.nsub;
  if (st.nguard) st.nguard[i] += s.nguard;
}

// The grid comes as a device copy.  The plain walk reads it through a plain
// pointer: its fields are then loaded where used instead of all occupying
// scalar registers from kernel entry (fewer spills).  The tracking walk
// (displacement record, per-step moments) stores to global memory in its
// loop, after which loads through a plain pointer would be repeated; it reads
// through a __restrict__ pointer, whose loads the compiler may keep.
__global__ void __launch_bounds__(256) NZ_WALK_ATTR walk_kernel(const WalkArgs A, const Grid* gp) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, A.xcd_chunk);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  const StepParams P{A.seed,  (uint32_t)A.base, (uint32_t)A.target, (uint32_t)A.limit,
                     (uint32_t)A.mom_base, A.dt, A.gph, A.garea, A.stepmom, A.flags, A.rec,
                     A.rec_ld};
  walk_slot<kPlain>(*gp, A.st, A.n, i, i < A.n, P, nullptr);
}

__global__ void __launch_bounds__(256) NZ_TRACK_ATTR moments_kernel(const WalkArgs A,
                                                                  const Grid* __restrict__ gpr) {
  // per-step moment partials of the block (12 KiB at kMomChunk = 256)
  __shared__ double smom[kMomChunk * 6];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, A.xcd_chunk);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  const StepParams P{A.seed,  (uint32_t)A.base, (uint32_t)A.target, (uint32_t)A.limit,
                     (uint32_t)A.mom_base, A.dt, A.gph, A.garea, A.stepmom, A.flags, A.rec,
                     A.rec_ld};
  extern __shared__ double mom_xs[];  // 4 per thread: start position
  walk_slot<kMoments>(*gpr, A.st, A.n, i, i < A.n, P, smom, mom_xs);
}

__global__ void __launch_bounds__(256) NZ_TRACK_ATTR record_kernel(const WalkArgs A,
                                                                 const Grid* __restrict__ gpr) {
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, A.xcd_chunk);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  const StepParams P{A.seed,  (uint32_t)A.base, (uint32_t)A.target, (uint32_t)A.limit,
                     (uint32_t)A.mom_base, A.dt, A.gph, A.garea, nullptr, A.flags, A.rec,
                     A.rec_ld, (uint32_t)A.rec_ring, (uint32_t)A.rec_pos0};
  extern __shared__ double rec_xs[];  // 4 per thread: start position, record row offset
  walk_slot<kRecord>(*gpr, A.st, A.n, i, i < A.n, P, nullptr, rec_xs);
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
    st.cell[a * n + i] = s.l[a];
    st.wrap[a * n + i] = 0;
    st.x[a * n + i] = s.x[a];
    st.x0[a * n + i] = s.x[a];
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

// In-box cell of a slot (its local cell index, the halo folded onto the
// nearest in-box cell: a spin exactly on the x = +L/2 face, or the rare one a
// step carried beyond the box).
__device__ __forceinline__ int32_t box_cell(const Grid& g, const int32_t* cell, int64_t n,
                                            int64_t i) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int l = cell[a * n + i] - 1;
    c[a] = l < 0 ? 0 : (l >= g.ncell[a] ? g.ncell[a] - 1 : l);
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
  // intra-structure spins first, then extra-structure spins; within a class
  // the cells' sort ranks (regions, then list length): spins with similar
  // collision rates and candidate counts share wavefronts
  const int64_t ncells = (int64_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
  const int32_t k = valid ? g.cell_rank[box_cell(g, st.cell, n, i)] +
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

// With nkeys_dev the scan length is *nkeys_dev + 1 (a batch's key space,
// which changes with its membership) and surplus blocks return.
__device__ __forceinline__ int64_t scan_len(int64_t m, const int64_t* nkeys_dev) {
  return nkeys_dev ? *nkeys_dev + 1 : m;
}

__global__ void __launch_bounds__(kScanThreads) tile_sum_kernel(const int32_t* counts, int64_t m,
                                                                int32_t* partial,
                                                                const int64_t* nkeys_dev) {
  m = scan_len(m, nkeys_dev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  if (base >= m) return;
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
__global__ void __launch_bounds__(kScanThreads) partial_scan_kernel(int32_t* partial, int64_t nb,
                                                                    const int64_t* nkeys_dev) {
  if (nkeys_dev) nb = (*nkeys_dev + 1 + kScanTile - 1) / kScanTile;
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
                                                                 const int32_t* partial,
                                                                 const int64_t* nkeys_dev) {
  __shared__ int32_t tile[kScanTile + kScanTile / 32];  // +1 word per 32: fewer bank conflicts
  m = scan_len(m, nkeys_dev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  if (base >= m) return;
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

int64_t scan_partial_len(int64_t m) { return (m + kScanTile - 1) / kScanTile; }

int exclusive_scan_i32(int32_t* counts, int64_t m, int32_t* partial, hipStream_t stream) {
  const int64_t nb = scan_partial_len(m);
  if (nb <= 0) return 0;
  hipLaunchKernelGGL(tile_sum_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, counts, m,
                     partial, (const int64_t*)nullptr);
  hipLaunchKernelGGL(partial_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, partial, nb,
                     (const int64_t*)nullptr);
  hipLaunchKernelGGL(tile_scan_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, stream, counts,
                     m, (const int32_t*)partial, (const int64_t*)nullptr);
  NZ_HIP(hipGetLastError());
  return 0;
}

// Slots [0, n) of src (fields strided by ld) to their sorted places in dst.
__global__ void __launch_bounds__(256) sort_scatter_kernel(const SlotState src, const SlotState dst,
                                                           int64_t n, int64_t ld,
                                                           const int32_t* keys,
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
    dst.cell[a * ld + j] = src.cell[a * ld + i];
    dst.x[a * ld + j] = src.x[a * ld + i];
    dst.wrap[a * ld + j] = src.wrap[a * ld + i];
    dst.x0[a * ld + j] = src.x0[a * ld + i];
    dst.mom[a * ld + j] = src.mom[a * ld + i];
  }
  dst.sid[j] = src.sid[i];
  dst.logsig[j] = src.logsig[i];
  dst.tstep[j] = src.tstep[i];
  if (src.nsub) dst.nsub[j] = src.nsub[i];
  if (src.nguard) dst.nguard[j] = src.nguard[i];
}

// ---- read-back in spin order ---------------------------------------------
__global__ void __launch_bounds__(256) read_kernel(const Grid g, const SlotState st, int64_t ld,
                                                   int64_t nslots, int64_t n, float* pos,
                                                   float* pos0, float* sig, float* mom,
                                                   double* pos_d, double* disp_d, int32_t* nsub,
                                                   int32_t* nguard) {
  // slots [0, nslots) of st (fields strided by ld); outputs [.., n] in spin order
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  const int64_t k = st.spin_id[i];
  if (k < 0 || k >= n) return;  // padding slots hold -1
  Lane s;
  double x0[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.x[a] = st.x[a * ld + i];
    s.l[a] = st.cell[a * ld + i];
    s.w[a] = st.wrap[a * ld + i];
    x0[a] = st.x0[a * ld + i];
  }
  if (pos || pos0) {
    float xf[3], x0f[3];
    lane_outputs(g, s, x0, xf, x0f);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (pos) pos[a * n + k] = xf[a];
      if (pos0) pos0[a * n + k] = x0f[a];
    }
  }
  if (sig) sig[k] = expf(st.logsig[i]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (mom) mom[a * n + k] = st.mom[a * ld + i];
    if (pos_d) pos_d[a * n + k] = s.x[a];
    if (disp_d) disp_d[a * n + k] = lane_disp(g, s, x0, a);
  }
  if (nsub) nsub[k] = st.nsub ? st.nsub[i] : 0;
  if (nguard) nguard[k] = st.nguard ? st.nguard[i] : 0;
}

__global__ void __launch_bounds__(256) disp_moments_kernel(const Grid g, const SlotState st,
                                                           int64_t n, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
    Lane s;
    double x0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s.x[a] = st.x[a * n + i];
      s.w[a] = st.wrap[a * n + i];
      x0[a] = st.x0[a * n + i];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double d = lane_disp(g, s, x0, a);
      v[a] = d * d;
      v[3 + a] = v[a] * v[a];
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double s = wave_sum_d(v[q]);
    if ((threadIdx.x & 63) == 0) atomicAdd(&out[q], s);
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
                     w->partial, (const int64_t*)nullptr);
  hipLaunchKernelGGL(partial_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, w->partial, nb,
                     (const int64_t*)nullptr);
  hipLaunchKernelGGL(tile_scan_kernel, dim3(nb), dim3(kScanThreads), 0, stream, w->counts, m,
                     w->partial, (const int64_t*)nullptr);
  hipLaunchKernelGGL(sort_scatter_kernel, dim3(blocks_for(n)), dim3(256), 0, stream,
                     w->buf[w->cur], w->buf[1 - w->cur], n, n, w->keys, w->counts, w->cursor);
  NZ_HIP(hipGetLastError());
  w->cur = 1 - w->cur;
  w->since_sort = 0;
  return 0;
}

static int check_walker(const Walker* w, const char* who) {
  if (!w) return fail(-1, std::string(who) + ": walker is NULL");
  return 0;
}

// ---- many simulations in one launch ---------------------------------------
// A batch holds a pool of slots (256 per block) shared by up to max_sims
// simulations, each with its own substrate, spin count, Philox key, time step,
// length and phase coefficients.  A simulation owns a contiguous range of
// whole blocks, so every workgroup walks exactly one simulation and reads its
// descriptor with uniform loads; its padding slots hold spin id -1.  One
// launch advances every active simulation by up to k steps; spatial sorting
// keeps each simulation's spins inside its range (the key space is laid out
// in pool order: a simulation's class/cell keys and one padding key, one key
// per free gap), so the walk kernel, sort and scan are those of the single
// walker.  The launch sequence (sort, walk, advance) depends on no host value
// that membership changes, so it is captured once in a HIP graph and replayed.

constexpr int64_t kBlk = 256;

struct SimDesc {
  Grid g;
  uint64_t seed;
  double dt;
  const float* gph;
  const float* garea;
  uint32_t flags;
  int32_t active;   // 1 while the simulation walks
  int64_t nsteps;   // walk length
  int64_t step;     // steps every spin of the simulation has completed
  int64_t k;        // steps per launch
  int64_t key0;     // first sort key
  int64_t ncells;   // cells of its grid
};

struct BatchJob {
  bool used = false;
  const Substrate* sub = nullptr;
  int64_t n = 0, b0 = 0, nb = 0;
};

struct Batch {
  int64_t cap = 0, nblocks = 0;
  int64_t active_blocks = 0;   // blocks up to the last one a simulation owns
  int64_t launch_blocks = 64;  // blocks a launch covers (>= active_blocks)
  int32_t max_sims = 0;
  SlotState buf[2];
  int cur = 0;
  int32_t* keys = nullptr;
  int32_t* counts = nullptr;
  int32_t* cursor = nullptr;
  int32_t* partial = nullptr;
  int64_t max_keys = 0;        // allocated key space
  int64_t* nkeys_dev = nullptr;
  SimDesc* descs_dev = nullptr;
  int32_t* block_sim_dev = nullptr;
  int64_t* block_key_dev = nullptr;
  float* pos_tmp = nullptr;    // [3, cap] seeding output
  int32_t* nfailed_dev = nullptr;
  std::vector<SimDesc> descs;
  std::vector<int32_t> block_sim;
  std::vector<int64_t> block_key;
  std::vector<BatchJob> jobs;
  bool dirty = true;           // host tables newer than the device's
  int32_t xcd_chunk = 64;
  int32_t walk_block = 64;     // threads per batch_walk_kernel workgroup (NZ_WALK_BLOCK)
  hipStream_t stream = nullptr;  // own stream: launches and graph capture
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  int graph_cur = 0;           // slot buffer the captured pair of launches starts from
  std::vector<void*> allocations;
  std::vector<void*> key_allocations;
  int device = 0;
};

__global__ void fill_i32_dev(int32_t* p, const int64_t* nkeys_dev, int64_t extra, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *nkeys_dev + extra) p[i] = v;
}

// Slots [slot0, slot0 + nslots) of a simulation from its n start positions
// [3, n]; padding slots get spin id -1.
__global__ void __launch_bounds__(256) batch_init_kernel(const Grid g, const SlotState st,
                                                         int64_t ld, int64_t slot0, int64_t nslots,
                                                         int64_t n, const float* pos) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nslots) return;
  const int64_t i = slot0 + e;
  if (e >= n) {
    st.spin_id[i] = -1;
    st.tstep[i] = 0u;
    return;
  }
  Lane s;
  const float p[3] = {pos[e], pos[n + e], pos[2 * n + e]};
  lane_init(g, p, s);
  st.spin_id[i] = (int32_t)e;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.cell[a * ld + i] = s.l[a];
    st.wrap[a * ld + i] = 0;
    st.x[a * ld + i] = s.x[a];
    st.x0[a * ld + i] = s.x[a];
    st.mom[a * ld + i] = 0.0f;
  }
  st.sid[i] = (int32_t)s.sid;
  st.logsig[i] = 0.0f;
  st.tstep[i] = 0u;
}

__global__ void __launch_bounds__(256) batch_keys_kernel(const SimDesc* descs,
                                                         const int32_t* block_sim,
                                                         const int64_t* block_key,
                                                         const SlotState st, int64_t ld,
                                                         int64_t nslots, int32_t* keys,
                                                         int32_t* counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  const int64_t pb = i / kBlk;
  const int32_t s = block_sim[pb];
  int64_t key;
  if (s < 0) {
    key = block_key[pb];
  } else {
    const SimDesc& d = descs[s];
    const Grid& g = d.g;
    if (st.spin_id[i] < 0) {
      key = d.key0 + 2 * d.ncells;
    } else {
      const int64_t cell = box_cell(g, st.cell, ld, i);
      key = d.key0 + g.cell_rank[cell] + (st.sid[i] >= g.nstruct ? d.ncells : 0);
    }
  }
  keys[i] = (int32_t)key;
  const WaveRun r = wave_run((int32_t)key, true);
  if (r.head == (int)(threadIdx.x & 63)) atomicAdd(&counts[key], r.len);
}

// A workgroup of blockDim.x (64, 128 or 256) threads walks one part of a
// pool block; the XCD remap runs over workgroups, chunk pool blocks at a time.
__global__ void __launch_bounds__(256) NZ_WALK_ATTR batch_walk_kernel(const SimDesc* descs,
                                                                      const int32_t* block_sim,
                                                                      const SlotState st,
                                                                      int64_t ld, int32_t chunk) {
  const int parts = kBlk / blockDim.x;
  const int64_t wg = xcd_block(blockIdx.x, gridDim.x, (int64_t)chunk * parts);
  const int64_t pb = wg / parts;
  const int32_t s = block_sim[pb];
  if (s < 0) return;
  const SimDesc& d = descs[s];
  if (!d.active) return;
  const int64_t i = pb * kBlk + (wg % parts) * blockDim.x + threadIdx.x;
  const int64_t target = min(d.step + d.k, d.nsteps);
  const StepParams P{d.seed, 0u, (uint32_t)target, (uint32_t)d.nsteps, 0u, d.dt, d.gph, d.garea,
                     nullptr, d.flags, nullptr};
  walk_slot<kPlain>(d.g, st, ld, i, st.spin_id[i] >= 0, P, nullptr);
}

// After a launch: every active simulation has advanced to its target.
__global__ void batch_advance_kernel(SimDesc* descs, int32_t max_sims) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= max_sims) return;
  SimDesc& d = descs[s];
  if (!d.active) return;
  d.step = min(d.step + d.k, d.nsteps);
  if (d.step >= d.nsteps) d.active = 0;
}

static int batch_alloc(Batch* b, size_t bytes, void** out, std::vector<void*>& list) {
  void* p = nullptr;
  NZ_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16)));
  list.push_back(p);
  *out = p;
  return 0;
}

static void batch_release(Batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  for (void* p : b->key_allocations) (void)hipFree(p);
  b->key_allocations.clear();
  for (void* p : b->allocations) (void)hipFree(p);
  b->allocations.clear();
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->ev_in = b->ev_out = nullptr;
  b->stream = nullptr;
}

// Key space and per-block tables from the membership (pool order).
static int64_t batch_layout(Batch* b) {
  int64_t key = 0;
  int64_t pb = 0;
  b->active_blocks = 0;
  while (pb < b->nblocks) {
    const int32_t s = b->block_sim[pb];
    if (s >= 0) {
      SimDesc& d = b->descs[s];
      d.key0 = key;
      key += 2 * d.ncells + 1;
      const int64_t e = pb + b->jobs[s].nb;
      for (; pb < e; ++pb) b->block_key[pb] = d.key0;
      b->active_blocks = pb;
    } else {
      for (; pb < b->nblocks && b->block_sim[pb] < 0; ++pb) b->block_key[pb] = key;
      ++key;
    }
  }
  return key;
}

// Grow the key arrays to hold nkeys keys (drops the captured graph).
static int batch_ensure_keys(Batch* b, int64_t nkeys) {
  if (nkeys <= b->max_keys) return 0;
  if (nkeys > (int64_t)INT32_MAX - 1) return fail(-1, "batch: sort key space exceeds 2^31");
  NZ_HIP(hipStreamSynchronize(b->stream));
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  for (void* p : b->key_allocations) (void)hipFree(p);
  b->key_allocations.clear();
  const int64_t m = std::max<int64_t>(nkeys + nkeys / 2, 1 << 16);
  void* p;
  if (int rc = batch_alloc(b, sizeof(int32_t) * (m + 1), &p, b->key_allocations)) return rc;
  b->counts = (int32_t*)p;
  if (int rc = batch_alloc(b, sizeof(int32_t) * (m + 1), &p, b->key_allocations)) return rc;
  b->cursor = (int32_t*)p;
  if (int rc = batch_alloc(b, sizeof(int32_t) * ((m + 1 + kScanTile - 1) / kScanTile), &p,
                           b->key_allocations))
    return rc;
  b->partial = (int32_t*)p;
  b->max_keys = m;
  return 0;
}

// Blocks a launch covers: a power of two (at least 64) at or above the last
// block a simulation owns, grown as soon as a simulation lands beyond it and
// shrunk when the used range falls to a quarter of it.  A change drops the
// captured graph; callers hold the stream synchronized.
static void batch_fit_launch(Batch* b) {
  int64_t want = 64;
  while (want < b->active_blocks) want <<= 1;
  want = std::min(want, b->nblocks);
  if (b->active_blocks > b->launch_blocks || want * 4 <= b->launch_blocks) {
    if (want != b->launch_blocks && b->graph) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
    b->launch_blocks = want;
  }
}

// Bring the device's progress fields to the host mirror (before the host
// rewrites the tables) and upload the tables when they changed.
static int batch_sync_tables(Batch* b, bool upload) {
  NZ_HIP(hipStreamSynchronize(b->stream));
  std::vector<SimDesc> dev(b->max_sims);
  NZ_HIP(hipMemcpy(dev.data(), b->descs_dev, sizeof(SimDesc) * b->max_sims, hipMemcpyDeviceToHost));
  for (int32_t s = 0; s < b->max_sims; ++s) {
    b->descs[s].step = dev[s].step;
    b->descs[s].active = dev[s].active;
  }
  if (!upload) return 0;
  const int64_t nkeys = batch_layout(b);
  if (int rc = batch_ensure_keys(b, nkeys)) return rc;
  batch_fit_launch(b);
  NZ_HIP(hipMemcpy(b->descs_dev, b->descs.data(), sizeof(SimDesc) * b->max_sims,
                   hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(b->block_sim_dev, b->block_sim.data(), sizeof(int32_t) * b->nblocks,
                   hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(b->block_key_dev, b->block_key.data(), sizeof(int64_t) * b->nblocks,
                   hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(b->nkeys_dev, &nkeys, sizeof(int64_t), hipMemcpyHostToDevice));
  b->dirty = false;
  return 0;
}

// One launch: sort every simulation's spins, walk, advance.
static int batch_enqueue_launch(Batch* b) {
  hipStream_t s = b->stream;
  const SlotState& src = b->buf[b->cur];
  const SlotState& dst = b->buf[1 - b->cur];
  const int64_t mk = b->max_keys;
  hipLaunchKernelGGL(fill_i32_dev, dim3(blocks_for(mk + 1)), dim3(256), 0, s, b->counts,
                     (const int64_t*)b->nkeys_dev, (int64_t)1, 0);
  hipLaunchKernelGGL(fill_i32_dev, dim3(blocks_for(mk)), dim3(256), 0, s, b->cursor,
                     (const int64_t*)b->nkeys_dev, (int64_t)0, 0);
  // only the blocks up to the last used one: the pool's tail is free and
  // keeps its place (its keys come last in pool order)
  const int64_t nslots = b->launch_blocks * kBlk;
  hipLaunchKernelGGL(batch_keys_kernel, dim3(blocks_for(nslots)), dim3(256), 0, s,
                     (const SimDesc*)b->descs_dev, (const int32_t*)b->block_sim_dev,
                     (const int64_t*)b->block_key_dev, src, b->cap, nslots, b->keys, b->counts);
  const int64_t nbt = (mk + 1 + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(tile_sum_kernel, dim3(nbt), dim3(kScanThreads), 0, s, b->counts, mk + 1,
                     b->partial, (const int64_t*)b->nkeys_dev);
  hipLaunchKernelGGL(partial_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, b->partial, nbt,
                     (const int64_t*)b->nkeys_dev);
  hipLaunchKernelGGL(tile_scan_kernel, dim3(nbt), dim3(kScanThreads), 0, s, b->counts, mk + 1,
                     (const int32_t*)b->partial, (const int64_t*)b->nkeys_dev);
  hipLaunchKernelGGL(sort_scatter_kernel, dim3(blocks_for(nslots)), dim3(256), 0, s, src, dst,
                     nslots, b->cap, (const int32_t*)b->keys, (const int32_t*)b->counts, b->cursor);
  hipLaunchKernelGGL(batch_walk_kernel, dim3((unsigned)(b->launch_blocks * (kBlk / b->walk_block))),
                     dim3(b->walk_block), 0, s,
                     (const SimDesc*)b->descs_dev, (const int32_t*)b->block_sim_dev, dst, b->cap,
                     b->xcd_chunk);
  hipLaunchKernelGGL(batch_advance_kernel, dim3(blocks_for(b->max_sims)), dim3(256), 0, s,
                     b->descs_dev, b->max_sims);
  NZ_HIP(hipGetLastError());
  b->cur = 1 - b->cur;
  return 0;
}

static int check_batch(const Batch* b, const char* who) {
  if (!b) return fail(-1, std::string(who) + ": batch is NULL");
  return 0;
}

static int check_sim(const Batch* b, int32_t sim, const char* who) {
  if (int rc = check_batch(b, who)) return rc;
  if (sim < 0 || sim >= b->max_sims || !b->jobs[sim].used)
    return fail(-1, std::string(who) + ": no such simulation");
  return 0;
}

// SlotState of a simulation's range (pointers offset to its first slot).
static SlotState batch_view(const Batch* b, int32_t sim) {
  SlotState v = b->buf[b->cur];
  const int64_t o = b->jobs[sim].b0 * kBlk;
  v.spin_id += o;
  v.cell += o;
  v.x += o;
  v.wrap += o;
  v.x0 += o;
  v.sid += o;
  v.logsig += o;
  v.mom += o;
  v.tstep += o;
  return v;
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
  if (const char* e = getenv("NZ_WALK_BLOCK")) {
    const int v = atoi(e);
    if (v == 64 || v == 128 || v == 256) w->walk_block = v;
  }
  if (const char* e = getenv("NZ_TRACK_BLOCK")) {
    const int v = atoi(e);
    if (v == 64 || v == 128 || v == 256) w->track_block = v;
  }
  int rc = 0;
  for (int b = 0; b < 2 && !rc; ++b) {
    SlotState& s = w->buf[b];
    void* p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * n, &p))) break;
    s.spin_id = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * 3 * n, &p))) break;
    s.cell = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(double) * 3 * n, &p))) break;
    s.x = (double*)p;
    if ((rc = walker_alloc(w, sizeof(int32_t) * 3 * n, &p))) break;
    s.wrap = (int32_t*)p;
    if ((rc = walker_alloc(w, sizeof(double) * 3 * n, &p))) break;
    s.x0 = (double*)p;
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
  w->rec_ahead = 0;
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
  w->rec_ahead = 0;
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
    a.rec = nullptr;
    a.rec_ld = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (w->timing) {
      if (int rc = next_event_pair(w, &e0, &e1)) return rc;
      NZ_HIP(hipEventRecord(e0, s));
    }
    if (stepmom)
      hipLaunchKernelGGL(moments_kernel, dim3(blocks_for(w->n, w->track_block)), dim3(w->track_block),
                         (size_t)w->track_block * 4 * sizeof(double), s, a,
                         (const Grid*)w->sub->g_dev);
    else
      hipLaunchKernelGGL(walk_kernel, dim3(blocks_for(w->n, w->walk_block)), dim3(w->walk_block), 0,
                         s, a, (const Grid*)w->sub->g_dev);
    NZ_HIP(hipGetLastError());
    if (w->timing) NZ_HIP(hipEventRecord(e1, s));
    target += chunk;
    w->since_sort += (int32_t)chunk;
  }
  w->step = end;
  w->rec_ahead = 0;  // steps walked past a record block are not recorded by this walk
  return 0;
}

int nz_walk_record_ring(nz_walker_t ww, uint64_t seed, int32_t nsteps, int32_t ahead, double dt,
                        float* rec, int64_t ld, int32_t ring, int32_t pos0, int32_t* slot_spin,
                        void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walk_record")) return rc;
  if (nsteps <= 0 || nsteps > 4096) return fail(-1, "nz_walk_record: nsteps must be in [1, 4096]");
  if (ahead < 0) return fail(-1, "nz_walk_record: ahead must be >= 0");
  if (!(dt >= 0.0)) return fail(-1, "nz_walk_record: dt must be >= 0");
  if (!rec || !slot_spin) return fail(-1, "nz_walk_record: NULL rec or slot_spin");
  if (ring < 1 || ld < 3 * (int64_t)ring)
    return fail(-1, "nz_walk_record: ld must be at least 3 * ring, ring >= 1");
  if (pos0 < 0 || (int64_t)pos0 + nsteps > ring)
    return fail(-1, "nz_walk_record: the block must fit the ring from pos0 without wrapping");
  if ((int64_t)nsteps + ahead > ring)
    return fail(-1, "nz_walk_record: nsteps + ahead must not exceed the ring");
  if (w->step + nsteps + ahead > (int64_t)UINT32_MAX)
    return fail(-1, "nz_walk_record: step counter overflow");
  if (w->rec_ahead > 0 && (rec != w->rec_ptr || ld != w->rec_ld || ring != w->ring ||
                           pos0 != w->ring_next))
    return fail(-1, "nz_walk_record: the previous record block ran ahead into its ring: "
                    "continue in the same ring (rec, ld, ring) at entry " +
                        std::to_string(w->ring_next));
  if (w->rec_ahead > nsteps)
    return fail(-1, "nz_walk_record: the previous record block ran further ahead than this "
                    "block is long");
  hipStream_t s = (hipStream_t)stream;
  // spins stand at different steps after a block that ran ahead, and their
  // steps past it sit in its slots: no re-sort until a block ends with every
  // spin at its end
  const bool sort_due = w->sort_interval > 0 && w->since_sort >= w->sort_interval;
  if (sort_due && w->rec_ahead == 0)
    if (int rc = sort_spins(w, s)) return rc;
  // the block before a re-sort does not run ahead
  const bool sort_next = w->sort_interval > 0 && w->since_sort + nsteps >= w->sort_interval;
  if (sort_next) ahead = 0;
  WalkArgs a;
  a.st = w->buf[w->cur];
  a.n = w->n;
  a.seed = seed;
  a.base = w->step;
  a.mom_base = w->step;
  a.target = w->step + nsteps;
  // a spin whose block is done walks on, up to ahead steps, while others of
  // its wavefront finish theirs; the record is complete up to the block end
  a.limit = a.target + ahead;
  a.dt = dt;
  a.gph = nullptr;
  a.garea = nullptr;
  a.stepmom = nullptr;
  a.flags = 0;
  a.xcd_chunk = w->xcd_chunk;
  a.rec = rec;
  a.rec_ld = ld;
  a.rec_ring = ring;
  a.rec_pos0 = pos0;
  // one wave per workgroup, as the plain walk (precision form, 4e6 spins,
  // 1280 steps at K = 64: 0.531 s against 0.537 s in 256-thread workgroups)
  hipLaunchKernelGGL(record_kernel, dim3(blocks_for(w->n, w->walk_block)), dim3(w->walk_block),
                     (size_t)w->walk_block * 4 * sizeof(double), s, a, (const Grid*)w->sub->g_dev);
  NZ_HIP(hipGetLastError());
  NZ_HIP(hipMemcpyAsync(slot_spin, w->buf[w->cur].spin_id, sizeof(int32_t) * w->n,
                        hipMemcpyDeviceToDevice, s));
  w->step += nsteps;
  w->since_sort += nsteps;
  w->rec_ahead = ahead;
  w->ring_next = (pos0 + nsteps) % ring;
  w->ring = ring;
  w->rec_ld = ld;
  w->rec_ptr = rec;
  return 0;
}

int nz_walk_record(nz_walker_t ww, uint64_t seed, int32_t nsteps, double dt, float* rec,
                   int64_t ld, int32_t* slot_spin, void* stream) {
  if (ww && ((Walker*)ww)->rec_ahead > 0)
    return fail(-1, "nz_walk_record: the previous record block ran ahead: continue it with "
                    "nz_walk_record_ring");
  if (ld < 3 * (int64_t)nsteps || ld / 3 > INT32_MAX)
    return fail(-1, "nz_walk_record: ld must be at least 3 * nsteps");
  return nz_walk_record_ring(ww, seed, nsteps, 0, dt, rec, ld, (int32_t)(ld / 3), 0, slot_spin,
                             stream);
}

int nz_walker_read(nz_walker_t ww, float* pos, float* pos0, float* sig, float* mom, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_read")) return rc;
  hipLaunchKernelGGL(read_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[w->cur], w->n, w->n, w->n, pos, pos0, sig, mom, (double*)nullptr,
                     (double*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_walker_read_f64(nz_walker_t ww, double* pos, double* disp, void* stream) {
  Walker* w = (Walker*)ww;
  if (int rc = check_walker(w, "nz_walker_read_f64")) return rc;
  hipLaunchKernelGGL(read_kernel, dim3(blocks_for(w->n)), dim3(256), 0, (hipStream_t)stream,
                     w->sub->g, w->buf[w->cur], w->n, w->n, w->n, (float*)nullptr, (float*)nullptr,
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
                     w->sub->g, w->buf[w->cur], w->n, w->n, w->n, (float*)nullptr, (float*)nullptr,
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


int nz_batch_create(int64_t capacity, int32_t max_sims, nz_batch_t* out) {
  if (!out) return fail(-1, "nz_batch_create: out is NULL");
  *out = nullptr;
  if (capacity <= 0 || capacity > (int64_t)INT32_MAX - kBlk)
    return fail(-1, "nz_batch_create: capacity must be in [1, 2^31 - 256)");
  if (max_sims <= 0 || max_sims > 65536) return fail(-1, "nz_batch_create: max_sims must be in [1, 65536]");
  Batch* b = new Batch();
  b->nblocks = (capacity + kBlk - 1) / kBlk;
  b->cap = b->nblocks * kBlk;
  b->launch_blocks = std::min<int64_t>(64, b->nblocks);
  b->max_sims = max_sims;
  (void)hipGetDevice(&b->device);
  if (const char* e = getenv("NZ_WALK_XCD_CHUNK")) b->xcd_chunk = std::max(0, atoi(e));
  if (const char* e = getenv("NZ_WALK_BLOCK")) {
    const int v = atoi(e);
    if (v == 64 || v == 128 || v == 256) b->walk_block = v;
  }
  b->descs.assign(max_sims, SimDesc{});
  b->jobs.assign(max_sims, BatchJob{});
  b->block_sim.assign(b->nblocks, -1);
  b->block_key.assign(b->nblocks, 0);
  const int64_t n = b->cap;
  int rc = 0;
  void* p = nullptr;
  auto A = [&](size_t bytes) { return batch_alloc(b, bytes, &p, b->allocations); };
  for (int k = 0; k < 2 && !rc; ++k) {
    SlotState& s = b->buf[k];
    if ((rc = A(sizeof(int32_t) * n))) break;
    s.spin_id = (int32_t*)p;
    if ((rc = A(sizeof(int32_t) * 3 * n))) break;
    s.cell = (int32_t*)p;
    if ((rc = A(sizeof(double) * 3 * n))) break;
    s.x = (double*)p;
    if ((rc = A(sizeof(int32_t) * 3 * n))) break;
    s.wrap = (int32_t*)p;
    if ((rc = A(sizeof(double) * 3 * n))) break;
    s.x0 = (double*)p;
    if ((rc = A(sizeof(int32_t) * n))) break;
    s.sid = (int32_t*)p;
    if ((rc = A(sizeof(float) * n))) break;
    s.logsig = (float*)p;
    if ((rc = A(sizeof(float) * 3 * n))) break;
    s.mom = (float*)p;
    if ((rc = A(sizeof(uint32_t) * n))) break;
    s.tstep = (uint32_t*)p;
    s.nsub = s.nguard = nullptr;
  }
  if (!rc && !(rc = A(sizeof(int32_t) * n))) b->keys = (int32_t*)p;
  if (!rc && !(rc = A(sizeof(float) * 3 * n))) b->pos_tmp = (float*)p;
  if (!rc && !(rc = A(sizeof(int64_t)))) b->nkeys_dev = (int64_t*)p;
  if (!rc && !(rc = A(sizeof(int32_t)))) b->nfailed_dev = (int32_t*)p;
  if (!rc && !(rc = A(sizeof(SimDesc) * max_sims))) b->descs_dev = (SimDesc*)p;
  if (!rc && !(rc = A(sizeof(int32_t) * b->nblocks))) b->block_sim_dev = (int32_t*)p;
  if (!rc && !(rc = A(sizeof(int64_t) * b->nblocks))) b->block_key_dev = (int64_t*)p;
  if (!rc) {
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
    if (e != hipSuccess) rc = hip_fail(e, "nz_batch_create: stream/events");
  }
  // every slot starts as free padding
  for (int k = 0; k < 2 && !rc; ++k) {
    hipLaunchKernelGGL(fill_i32, dim3(blocks_for(n)), dim3(256), 0, b->stream, b->buf[k].spin_id, n,
                       -1);
    hipLaunchKernelGGL(fill_u32, dim3(blocks_for(n)), dim3(256), 0, b->stream, b->buf[k].tstep, n, 0u);
  }
  if (!rc) {
    const hipError_t e = hipMemcpy(b->descs_dev, b->descs.data(), sizeof(SimDesc) * max_sims,
                                   hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "nz_batch_create: descriptors");
  }
  if (!rc) rc = batch_sync_tables(b, true);
  if (rc) {
    batch_release(b);
    delete b;
    return rc;
  }
  *out = (nz_batch_t)b;
  return 0;
}

int nz_batch_destroy(nz_batch_t bb) {
  if (!bb) return 0;
  Batch* b = (Batch*)bb;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  (void)hipSetDevice(dev);
  delete b;
  return 0;
}

int nz_batch_add(nz_batch_t bb, nz_substrate_t sub, int64_t n, uint64_t seed, int64_t nsteps,
                 double dt, const float* g, const float* garea, uint32_t flags, int32_t k,
                 const float* pos, const int32_t* targets_host, int32_t ntargets, int32_t* sim_out,
                 void* stream) {
  Batch* b = (Batch*)bb;
  if (int rc = check_batch(b, "nz_batch_add")) return rc;
  if (!sub || !sim_out) return fail(-1, "nz_batch_add: NULL substrate or sim_out");
  if (n <= 0 || nsteps < 0 || nsteps > (int64_t)UINT32_MAX || !(dt >= 0.0) || k <= 0)
    return fail(-1, "nz_batch_add: need n > 0, 0 <= nsteps < 2^32, dt >= 0, k > 0");
  if ((g == nullptr) != (garea == nullptr)) return fail(-1, "nz_batch_add: pass both g and garea or neither");
  if (!pos && (!targets_host || ntargets <= 0)) return fail(-1, "nz_batch_add: need pos or targets");
  const Substrate* S = (const Substrate*)sub;
  const int64_t nb = (n + kBlk - 1) / kBlk;
  int32_t sim = -1;
  for (int32_t s = 0; s < b->max_sims; ++s)
    if (!b->jobs[s].used) { sim = s; break; }
  int64_t b0 = -1;
  for (int64_t pb = 0, run = 0; pb < b->nblocks; ++pb) {  // first fit
    run = (b->block_sim[pb] < 0) ? run + 1 : 0;
    if (run == nb) { b0 = pb - nb + 1; break; }
  }
  if (sim < 0 || b0 < 0) return fail(NZ_EFULL, "nz_batch_add: no room (simulation slots or contiguous spin blocks)");
  if (int rc = batch_sync_tables(b, false)) return rc;
  // order the caller's stream work (pos, g) before ours
  NZ_HIP(hipEventRecord(b->ev_in, (hipStream_t)stream));
  NZ_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  const float* p0 = pos;
  if (!p0) {
    hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, b->stream, b->nfailed_dev, (int64_t)1, 0);
    if (int rc = nz_seed(sub, n, targets_host, ntargets, seed ^ 0x5EED5EED5EEDull, kSeedAttempts,
                         b->pos_tmp, b->nfailed_dev, b->stream))
      return rc;
    int32_t nf = 0;
    NZ_HIP(hipMemcpyAsync(&nf, b->nfailed_dev, sizeof(nf), hipMemcpyDeviceToHost, b->stream));
    NZ_HIP(hipStreamSynchronize(b->stream));
    if (nf) return fail(-1, "nz_batch_add: " + std::to_string(nf) + " spins could not be seeded "
                            "(target structures empty or too small)");
    p0 = b->pos_tmp;
  }
  const int64_t slot0 = b0 * kBlk;
  for (int k2 = 0; k2 < 2; ++k2)  // both buffers: the range holds this simulation whichever is current
    hipLaunchKernelGGL(batch_init_kernel, dim3(blocks_for(nb * kBlk)), dim3(256), 0, b->stream,
                       S->g, b->buf[k2], b->cap, slot0, nb * kBlk, n, p0);
  NZ_HIP(hipGetLastError());
  SimDesc& d = b->descs[sim];
  d = SimDesc{};
  d.g = S->g;
  d.seed = seed;
  d.dt = dt;
  d.gph = g;
  d.garea = garea;
  d.flags = flags;
  d.active = nsteps > 0 ? 1 : 0;
  d.nsteps = nsteps;
  d.step = 0;
  d.k = k;
  d.ncells = (int64_t)S->g.ncell[0] * S->g.ncell[1] * S->g.ncell[2];
  b->jobs[sim] = BatchJob{true, S, n, b0, nb};
  for (int64_t pb = b0; pb < b0 + nb; ++pb) b->block_sim[pb] = sim;
  // the progress fields of the new entry must not be overwritten by the
  // device's stale copy: upload now
  const int64_t nkeys = batch_layout(b);
  if (int rc = batch_ensure_keys(b, nkeys)) return rc;
  NZ_HIP(hipStreamSynchronize(b->stream));
  batch_fit_launch(b);
  NZ_HIP(hipMemcpy(b->descs_dev, b->descs.data(), sizeof(SimDesc) * b->max_sims, hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(b->block_sim_dev, b->block_sim.data(), sizeof(int32_t) * b->nblocks,
                   hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(b->block_key_dev, b->block_key.data(), sizeof(int64_t) * b->nblocks,
                   hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(b->nkeys_dev, &nkeys, sizeof(int64_t), hipMemcpyHostToDevice));
  b->dirty = false;
  *sim_out = sim;
  return 0;
}

int nz_batch_remove(nz_batch_t bb, int32_t sim) {
  Batch* b = (Batch*)bb;
  if (int rc = check_sim(b, sim, "nz_batch_remove")) return rc;
  if (int rc = batch_sync_tables(b, false)) return rc;
  const BatchJob j = b->jobs[sim];
  for (int64_t pb = j.b0; pb < j.b0 + j.nb; ++pb) b->block_sim[pb] = -1;
  b->jobs[sim] = BatchJob{};
  b->descs[sim] = SimDesc{};
  // its slots become free padding in both buffers
  for (int k2 = 0; k2 < 2; ++k2)
    hipLaunchKernelGGL(fill_i32, dim3(blocks_for(j.nb * kBlk)), dim3(256), 0, b->stream,
                       b->buf[k2].spin_id + j.b0 * kBlk, j.nb * kBlk, -1);
  NZ_HIP(hipGetLastError());
  return batch_sync_tables(b, true);
}

int nz_batch_run(nz_batch_t bb, int32_t launches, void* stream) {
  Batch* b = (Batch*)bb;
  if (int rc = check_batch(b, "nz_batch_run")) return rc;
  if (launches < 0) return fail(-1, "nz_batch_run: launches < 0");
  if (b->dirty)
    if (int rc = batch_sync_tables(b, true)) return rc;
  NZ_HIP(hipEventRecord(b->ev_in, (hipStream_t)stream));
  NZ_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  int32_t left = launches;
  if (b->use_graph && left >= 2) {
    if (b->graph && b->graph_cur != b->cur) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
    if (!b->graph) {  // two launches: the slot buffers are back where they started
      b->graph_cur = b->cur;
      const int cur0 = b->cur;
      NZ_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = batch_enqueue_launch(b);
      if (!rc) rc = batch_enqueue_launch(b);
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &gr);
      b->cur = cur0;
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      NZ_HIP(e);
      const hipError_t e2 = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      NZ_HIP(e2);
    }
    for (; left >= 2; left -= 2) NZ_HIP(hipGraphLaunch(b->graph, b->stream));
  }
  for (; left > 0; --left)
    if (int rc = batch_enqueue_launch(b)) return rc;
  NZ_HIP(hipEventRecord(b->ev_out, b->stream));
  NZ_HIP(hipStreamWaitEvent((hipStream_t)stream, b->ev_out, 0));
  return 0;
}

int nz_batch_set_graph(nz_batch_t bb, int32_t enable) {
  Batch* b = (Batch*)bb;
  if (int rc = check_batch(b, "nz_batch_set_graph")) return rc;
  b->use_graph = enable != 0;
  return 0;
}

int nz_batch_status(nz_batch_t bb, int32_t sim, int64_t* step, int32_t* done) {
  Batch* b = (Batch*)bb;
  if (int rc = check_sim(b, sim, "nz_batch_status")) return rc;
  if (int rc = batch_sync_tables(b, false)) return rc;
  if (step) *step = b->descs[sim].step;
  if (done) *done = (b->descs[sim].active == 0) ? 1 : 0;
  return 0;
}

int nz_batch_status_all(nz_batch_t bb, int64_t* steps, int32_t* done) {
  Batch* b = (Batch*)bb;
  if (int rc = check_batch(b, "nz_batch_status_all")) return rc;
  if (int rc = batch_sync_tables(b, false)) return rc;
  for (int32_t s = 0; s < b->max_sims; ++s) {
    const bool used = b->jobs[s].used;
    if (steps) steps[s] = used ? b->descs[s].step : -1;
    if (done) done[s] = used ? (b->descs[s].active == 0 ? 1 : 0) : -1;
  }
  return 0;
}

int nz_batch_free(nz_batch_t bb, int64_t* largest_run, int64_t* free_slots, int32_t* free_sims) {
  Batch* b = (Batch*)bb;
  if (int rc = check_batch(b, "nz_batch_free")) return rc;
  int64_t best = 0, run = 0, total = 0;
  for (int64_t pb = 0; pb < b->nblocks; ++pb) {
    run = (b->block_sim[pb] < 0) ? run + 1 : 0;
    total += (b->block_sim[pb] < 0);
    best = std::max(best, run);
  }
  int32_t fs = 0;
  for (const BatchJob& j : b->jobs) fs += !j.used;
  if (largest_run) *largest_run = best * kBlk;
  if (free_slots) *free_slots = total * kBlk;
  if (free_sims) *free_sims = fs;
  return 0;
}

int nz_batch_read(nz_batch_t bb, int32_t sim, float* pos, float* pos0, float* sig, float* mom,
                  void* stream) {
  Batch* b = (Batch*)bb;
  if (int rc = check_sim(b, sim, "nz_batch_read")) return rc;
  const BatchJob& j = b->jobs[sim];
  NZ_HIP(hipEventRecord(b->ev_out, b->stream));
  NZ_HIP(hipStreamWaitEvent((hipStream_t)stream, b->ev_out, 0));
  hipLaunchKernelGGL(read_kernel, dim3(blocks_for(j.nb * kBlk)), dim3(256), 0, (hipStream_t)stream,
                     j.sub->g, batch_view(b, sim), b->cap, j.nb * kBlk, j.n, pos, pos0, sig, mom,
                     (double*)nullptr, (double*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_batch_signals(nz_batch_t bb, int32_t sim, int32_t W, const float* dirs, const float* scale,
                     double* out, void* stream) {
  Batch* b = (Batch*)bb;
  if (int rc = check_sim(b, sim, "nz_batch_signals")) return rc;
  const BatchJob& j = b->jobs[sim];
  NZ_HIP(hipEventRecord(b->ev_out, b->stream));
  NZ_HIP(hipStreamWaitEvent((hipStream_t)stream, b->ev_out, 0));
  const SlotState v = batch_view(b, sim);
  return signal_reduce_pooled(j.nb * kBlk, b->cap, v.spin_id, W, v.mom, dirs, scale, out,
                              (hipStream_t)stream);
}

}  // extern "C"
