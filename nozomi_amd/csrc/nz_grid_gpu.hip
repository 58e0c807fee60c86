// The cell grid of a substrate built on the device (nz_substrate_create).
//
// Replaces DiffSim3d.set_segments (diffsim3d.py:107-174), which builds its
// padded per-segment lists on the host in O(cells x spheres).  The lists,
// records and sort ranks equal those of the host builder (build_grid_host,
// nz_grid_build.h) entry for entry:
//   1. each sphere counts itself into the cells its AABB comes within the
//      list margin of (the host's axis ranges, in the same FP64 arithmetic);
//   2. an exclusive scan of the counts gives the CSR starts;
//   3. each sphere appends its index to its cells' lists (atomic cursors);
//   4. one thread per cell sorts its list by sphere index (the host appends
//      spheres in index order) and writes the cell-relative records;
//   5. one thread per in-box cell computes its sort rank: the cells of each
//      4x4x4 region in (list length, cell index) order, regions in index
//      order, the region's offset in closed form.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nz_grid_build.h"
#include "nz_internal.h"

// The cell ranges and record coordinates must round as the host builder's:
// no contraction of a*b+c into fma anywhere in this file.
#pragma clang fp contract(off)

namespace nz {

struct GridGeo {
  double L[3], hd[3];
  double M;          // list margin
  int32_t nloc[3];   // cells per axis with the halo
  int32_t ncell[3];  // cells per axis inside the box
  int64_t S;
  const double* sp;  // [5, S]: cx, cy, cz, r, sid
};

__device__ __forceinline__ double geo_cell_lo(const GridGeo& G, int a, int l) {
  return -G.L[a] / 2 + (l - 1) * G.hd[a];
}

// The host builder's axis_range (nz_grid_build.h) for sphere i on axis a.
__device__ __forceinline__ void geo_axis_range(const GridGeo& G, int64_t i, int a, int& lo,
                                               int& hi) {
  const double c = G.sp[(int64_t)a * G.S + i];
  const double r = G.sp[3 * G.S + i];
  lo = (int)floor((c - r - G.M + G.L[a] / 2) / G.hd[a]) - 1;
  hi = (int)floor((c + r + G.M + G.L[a] / 2) / G.hd[a]) + 2;
  lo = lo > 0 ? lo : 0;
  hi = hi < G.nloc[a] - 1 ? hi : G.nloc[a] - 1;
  while (lo <= hi && geo_cell_lo(G, a, lo) + G.hd[a] + G.M < c - r) ++lo;  // exact AABB test
  while (hi >= lo && geo_cell_lo(G, a, hi) - G.M > c + r) --hi;
}

__device__ __forceinline__ bool geo_range(const GridGeo& G, int64_t i, int lo[3], int hi[3]) {
  bool any = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    geo_axis_range(G, i, a, lo[a], hi[a]);
    any = any && lo[a] <= hi[a];
  }
  return any;
}

__global__ void grid_zero_kernel(int32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void grid_count_kernel(const GridGeo G, int32_t* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.S) return;
  int lo[3], hi[3];
  if (!geo_range(G, i, lo, hi)) return;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x)
        atomicAdd(&cnt[((int64_t)z * G.nloc[1] + y) * G.nloc[0] + x], 1);
}

// stats[0] = longest list, stats[1] = entries over all cells, stats[2] =
// entries over the in-box cells.
__global__ void grid_stats_kernel(const GridGeo G, const int32_t* cnt, int64_t ncells,
                                  unsigned long long* stats) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t n = cnt[c];
  if (n == 0) return;
  const int x = (int)(c % G.nloc[0]);
  const int y = (int)((c / G.nloc[0]) % G.nloc[1]);
  const int z = (int)(c / ((int64_t)G.nloc[0] * G.nloc[1]));
  atomicMax(&stats[0], (unsigned long long)n);
  atomicAdd(&stats[1], (unsigned long long)n);
  if (x >= 1 && x <= G.ncell[0] && y >= 1 && y <= G.ncell[1] && z >= 1 && z <= G.ncell[2])
    atomicAdd(&stats[2], (unsigned long long)n);
}

__global__ void grid_fill_kernel(const GridGeo G, int32_t* cursor, int32_t* ent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.S) return;
  int lo[3], hi[3];
  if (!geo_range(G, i, lo, hi)) return;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const int32_t k = atomicAdd(&cursor[((int64_t)z * G.nloc[1] + y) * G.nloc[0] + x], 1);
        ent[k] = (int32_t)i;
      }
}

// Padded length of each list in the reach array (even: lists start at even
// entries); the entry after the last cell is 0 for the scan.
__global__ void grid_reach_count_kernel(const int32_t* start, int64_t ncells, int32_t* rcnt) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  rcnt[c] = c < ncells ? ((start[c + 1] - start[c] + 1) & ~1) : 0;
}

__global__ void grid_cell_tab_kernel(const int32_t* start, const int32_t* reach_start,
                                     int64_t ncells, int4* tab) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) tab[c] = make_int4(start[c], start[c + 1], reach_start[c], 0);
}

// One thread per cell: sort the list by sphere index, then write the records
// and the reach entries (make_records, make_reach_entry: the host builder's
// arithmetic).
__global__ void grid_records_kernel(const GridGeo G, const int32_t* start, int64_t ncells,
                                    int32_t* ent, float4* cand, const int32_t* reach_start,
                                    uint2* reach, double lambda, double sigma) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t b = start[c], e = start[c + 1];
  for (int32_t k = b + 1; k < e; ++k) {  // insertion sort (lists are short)
    const int32_t v = ent[k];
    int32_t j = k - 1;
    while (j >= b && ent[j] > v) {
      ent[j + 1] = ent[j];
      --j;
    }
    ent[j + 1] = v;
  }
  const int x = (int)(c % G.nloc[0]);
  const int y = (int)((c / G.nloc[0]) % G.nloc[1]);
  const int z = (int)(c / ((int64_t)G.nloc[0] * G.nloc[1]));
  const double lx = geo_cell_lo(G, 0, x), ly = geo_cell_lo(G, 1, y), lz = geo_cell_lo(G, 2, z);
  const int32_t rs = reach_start[c];
  for (int32_t k = b; k < e; ++k) {
    const int32_t i = ent[k];
    const double cx = G.sp[i] - lx, cy = G.sp[G.S + i] - ly, cz = G.sp[2 * G.S + i] - lz;
    make_records(cx, cy, cz, G.sp[3 * G.S + i], G.sp[4 * G.S + i], i, cand[2 * (int64_t)k],
                 cand[2 * (int64_t)k + 1]);
    reach[rs + (k - b)] = make_reach_entry(cx, cy, cz, G.sp[3 * G.S + i], lambda, sigma);
  }
}

// Sort rank of in-box cell c = (z ny + y) nx + x: cells grouped in R^3
// regions (regions in index order ((z/R) ry + y/R) rx + x/R), by list length
// then cell index within a region.
__global__ void grid_rank_kernel(const GridGeo G, const int32_t* start, int32_t* rank) {
  constexpr int R = NZ_RANK_REGION;
  const int nx = G.ncell[0], ny = G.ncell[1], nz = G.ncell[2];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)nx * ny * nz) return;
  const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
  auto len = [&](int xx, int yy, int zz) {
    const int64_t lc = ((int64_t)(zz + 1) * G.nloc[1] + (yy + 1)) * G.nloc[0] + (xx + 1);
    return start[lc + 1] - start[lc];
  };
  const int qx = x / R, qy = y / R, qz = z / R;
  const int x0 = qx * R, y0 = qy * R, z0 = qz * R;
  const int sx = (nx - x0 < R) ? nx - x0 : R;
  const int sy = (ny - y0 < R) ? ny - y0 : R;
  const int sz = (nz - z0 < R) ? nz - z0 : R;
  // cells of the regions before this one: whole slabs below z0, whole rows of
  // this slab below y0, then the regions of this row before x0
  int64_t r = (int64_t)nx * ny * z0 + (int64_t)nx * y0 * sz + (int64_t)x0 * sy * sz;
  const int32_t n = len(x, y, z);
  for (int zz = z0; zz < z0 + sz; ++zz)
    for (int yy = y0; yy < y0 + sy; ++yy)
      for (int xx = x0; xx < x0 + sx; ++xx) {
        const int32_t m = len(xx, yy, zz);
        const int64_t cc = ((int64_t)zz * ny + yy) * nx + xx;
        r += (m < n || (m == n && cc < c)) ? 1 : 0;
      }
  rank[c] = (int32_t)r;
}

__global__ void grid_spheres_kernel(const GridGeo G, double4* sph, int32_t* sid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.S) return;
  sph[i] = make_double4(G.sp[i], G.sp[G.S + i], G.sp[2 * G.S + i], G.sp[3 * G.S + i]);
  sid[i] = (int32_t)G.sp[4 * G.S + i];
}

static unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

template <class T>
static int dev_alloc(Substrate* s, size_t count, T** out) {
  void* p = nullptr;
  const size_t bytes = count * sizeof(T) > 16 ? count * sizeof(T) : 16;
  NZ_HIP(hipMalloc(&p, bytes));
  s->allocations.push_back(p);
  s->info.device_bytes += (int64_t)bytes;
  *out = (T*)p;
  return 0;
}

// Scratch of the build, freed on every return path.
struct GridScratch {
  std::vector<void*> p;
  ~GridScratch() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int alloc(size_t count, T** out) {
    void* q = nullptr;
    NZ_HIP(hipMalloc(&q, count * sizeof(T) > 16 ? count * sizeof(T) : 16));
    p.push_back(q);
    *out = (T*)q;
    return 0;
  }
};

int build_grid_device(const double* spheres, int64_t S, int32_t nstruct, const double* D,
                      const double* T2, const double L[3], double cell_size, double margin,
                      Substrate* s) {
  Grid& g = s->g;
  double hd[3], M;
  std::string err;
  if (grid_params(spheres, S, nstruct, D, T2, L, cell_size, margin, g, hd, M, err) != 0)
    return fail(-1, "nz_substrate_create: " + err);
  std::memset(&s->info, 0, sizeof(s->info));
  const int64_t ncells = (int64_t)g.nloc[0] * g.nloc[1] * g.nloc[2];
  const int64_t nbox = (int64_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
  const hipStream_t st = nullptr;
  GridScratch tmp;
  GridGeo G;
  for (int a = 0; a < 3; ++a) {
    G.L[a] = L[a];
    G.hd[a] = hd[a];
    G.nloc[a] = g.nloc[a];
    G.ncell[a] = g.ncell[a];
  }
  G.M = M;
  G.S = S;
  double* sp = nullptr;
  if (int rc = tmp.alloc(5 * (size_t)S, &sp)) return rc;
  if (S > 0) NZ_HIP(hipMemcpy(sp, spheres, sizeof(double) * 5 * (size_t)S, hipMemcpyHostToDevice));
  G.sp = sp;

  // 1. counts (the last entry stays 0 for the scan) and their statistics
  int32_t *cnt = nullptr, *partial = nullptr, *cursor = nullptr;
  unsigned long long* stats = nullptr;
  if (int rc = tmp.alloc((size_t)ncells + 1, &cnt)) return rc;
  if (int rc = tmp.alloc((size_t)scan_partial_len(ncells + 1) + 1, &partial)) return rc;
  if (int rc = tmp.alloc((size_t)ncells, &cursor)) return rc;
  if (int rc = tmp.alloc(4, &stats)) return rc;
  hipLaunchKernelGGL(grid_zero_kernel, dim3(blocks256(ncells + 1)), dim3(256), 0, st, cnt,
                     ncells + 1);
  hipLaunchKernelGGL(grid_zero_kernel, dim3(1), dim3(256), 0, st, (int32_t*)stats, (int64_t)8);
  if (S > 0)
    hipLaunchKernelGGL(grid_count_kernel, dim3(blocks256(S)), dim3(256), 0, st, G, cnt);
  hipLaunchKernelGGL(grid_stats_kernel, dim3(blocks256(ncells)), dim3(256), 0, st, G,
                     (const int32_t*)cnt, ncells, stats);
  NZ_HIP(hipGetLastError());
  unsigned long long hs[4];
  NZ_HIP(hipMemcpy(hs, stats, sizeof(hs), hipMemcpyDeviceToHost));
  if (hs[1] > (unsigned long long)kMaxCandEntries)
    return fail(-1, "nz_substrate_create: candidate lists exceed 2^26 entries (2 GiB of records); "
                    "lower cell_size or margin");
  const int64_t total = (int64_t)hs[1];

  // 2. CSR starts (kept), 3. lists, 4. records
  int32_t *start = nullptr, *rank = nullptr, *ent = nullptr, *sid = nullptr;
  float4* cand = nullptr;
  double4* sph = nullptr;
  if (int rc = dev_alloc(s, (size_t)ncells + 1, &start)) return rc;
  if (int rc = dev_alloc(s, (size_t)nbox, &rank)) return rc;
  if (int rc = dev_alloc(s, 2 * ((size_t)total + kCandPad), &cand)) return rc;  // + padding
  if (int rc = dev_alloc(s, (size_t)S, &sph)) return rc;
  if (int rc = dev_alloc(s, (size_t)S, &sid)) return rc;
  if (int rc = tmp.alloc((size_t)total, &ent)) return rc;
  if (int rc = exclusive_scan_i32(cnt, ncells + 1, partial, st)) return rc;
  NZ_HIP(hipMemcpyAsync(start, cnt, sizeof(int32_t) * ((size_t)ncells + 1),
                        hipMemcpyDeviceToDevice, st));
  NZ_HIP(hipMemcpyAsync(cursor, cnt, sizeof(int32_t) * (size_t)ncells, hipMemcpyDeviceToDevice,
                        st));
  hipLaunchKernelGGL(grid_zero_kernel, dim3(1), dim3(256), 0, st,
                     (int32_t*)(cand + 2 * total), (int64_t)(8 * kCandPad));
  if (S > 0) {
    hipLaunchKernelGGL(grid_fill_kernel, dim3(blocks256(S)), dim3(256), 0, st, G, cursor, ent);
    hipLaunchKernelGGL(grid_spheres_kernel, dim3(blocks256(S)), dim3(256), 0, st, G, sph, sid);
  }
  // reach starts: scan of the padded list lengths (the partial buffer of the
  // first scan is free again: same stream)
  int32_t* reach_start = nullptr;
  if (int rc = dev_alloc(s, (size_t)ncells + 1, &reach_start)) return rc;
  hipLaunchKernelGGL(grid_reach_count_kernel, dim3(blocks256(ncells + 1)), dim3(256), 0, st,
                     (const int32_t*)start, ncells, reach_start);
  if (int rc = exclusive_scan_i32(reach_start, ncells + 1, partial, st)) return rc;
  int32_t rtotal = 0;
  NZ_HIP(hipMemcpyAsync(&rtotal, reach_start + ncells, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  NZ_HIP(hipStreamSynchronize(st));
  if (((int64_t)rtotal + kReachPad) * 8 >= ((int64_t)1 << 31))
    return fail(-1, "nz_substrate_create: reach entries exceed 2 GiB; lower cell_size or margin");
  uint2* reach = nullptr;
  if (int rc = dev_alloc(s, (size_t)rtotal + kReachPad, &reach)) return rc;
  int4* cell_tab = nullptr;
  if (int rc = dev_alloc(s, (size_t)ncells, &cell_tab)) return rc;
  hipLaunchKernelGGL(grid_cell_tab_kernel, dim3(blocks256(ncells)), dim3(256), 0, st,
                     (const int32_t*)start, (const int32_t*)reach_start, ncells, cell_tab);
  const int64_t rwords = 2 * ((int64_t)rtotal + kReachPad);
  hipLaunchKernelGGL(grid_zero_kernel, dim3(blocks256(rwords)), dim3(256), 0, st, (int32_t*)reach,
                     rwords);
  hipLaunchKernelGGL(grid_records_kernel, dim3(blocks256(ncells)), dim3(256), 0, st, G,
                     (const int32_t*)start, ncells, ent, cand, (const int32_t*)reach_start, reach,
                     g.reach_lambda, (double)g.reach_sigma);
  // 5. sort ranks
  hipLaunchKernelGGL(grid_rank_kernel, dim3(blocks256(nbox)), dim3(256), 0, st, G,
                     (const int32_t*)start, rank);
  NZ_HIP(hipGetLastError());

  // per-structure parameters (host side: a few hundred values)
  std::vector<double> Dv(D, D + nstruct + 1), T2v(T2, T2 + nstruct + 1);
  std::vector<float> invT2(nstruct + 1);
  for (int32_t k = 0; k <= nstruct; ++k) invT2[k] = (float)(1.0 / T2[k]);
  double *Dd = nullptr, *T2d = nullptr;
  float* invd = nullptr;
  if (int rc = dev_alloc(s, Dv.size(), &Dd)) return rc;
  if (int rc = dev_alloc(s, T2v.size(), &T2d)) return rc;
  if (int rc = dev_alloc(s, invT2.size(), &invd)) return rc;
  NZ_HIP(hipMemcpy(Dd, Dv.data(), sizeof(double) * Dv.size(), hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(T2d, T2v.data(), sizeof(double) * T2v.size(), hipMemcpyHostToDevice));
  NZ_HIP(hipMemcpy(invd, invT2.data(), sizeof(float) * invT2.size(), hipMemcpyHostToDevice));
  NZ_HIP(hipStreamSynchronize(st));

  g.cell_start = start;
  g.cell_rank = rank;
  g.cand = cand;
  g.cand_bytes = (int32_t)(sizeof(float4) * 2 * ((size_t)total + kCandPad));
  g.reach = reach;
  g.reach_start = reach_start;
  g.reach_bytes = (int32_t)(sizeof(uint2) * ((size_t)rtotal + kReachPad));
  g.cell_tab = cell_tab;
  g.sph = sph;
  g.sph_sid = sid;
  g.D = Dd;
  g.T2 = T2d;
  g.inv_T2 = invd;
  if (int rc = dev_alloc(s, 1, &s->g_dev)) return rc;
  NZ_HIP(hipMemcpy(s->g_dev, &g, sizeof(Grid), hipMemcpyHostToDevice));
  nz_substrate_info_t& inf = s->info;
  inf.nspheres = S;
  inf.nstructures = nstruct;
  for (int a = 0; a < 3; ++a) {
    inf.ncell[a] = g.ncell[a];
    inf.cell[a] = hd[a];
  }
  inf.margin = margin;
  inf.ncand_total = total;
  inf.ncand_records = total + kCandPad;
  inf.ncand_max = (int32_t)hs[0];
  inf.ncand_mean_box = (double)hs[2] / (double)nbox;
  return 0;
}

}  // namespace nz
