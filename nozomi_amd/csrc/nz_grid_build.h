// Host-side construction of the walker's cell grid (used by
// nz_substrate_create and by the CPU build of the walker in the tests).
//
// The box [-L/2, L/2] per axis is tiled exactly by ncell cells; one halo cell
// per face holds the explicit image spheres beyond it.  A cell's list keeps
// every sphere whose AABB comes within margin (+ slack for the tolerance band
// and coordinate rounding) of the cell box, with coordinates relative to the
// cell's lower corner.  Compare the reference's per-segment lists with a 1 um
// AABB margin (diffsim3d.py:107-174).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nz_internal.h"

// Edge (in cells) of the regions within which cells are ordered by list
// length for the spatial sort.
#ifndef NZ_RANK_REGION
#define NZ_RANK_REGION 4
#endif

namespace nz {

struct HostGrid {
  Grid g;  // scalar fields set; pointers bound by bind_host() or by the upload
  std::vector<int32_t> start;
  std::vector<int32_t> rank;
  std::vector<float4> cand;  // a, b records interleaved
  std::vector<double4> sph;
  std::vector<int32_t> sid;
  std::vector<double> D, T2;
  std::vector<float> inv_T2;
  nz_substrate_info_t info;

  void bind_host() {
    g.cell_start = start.data();
    g.cell_rank = rank.data();
    g.cand = cand.data();
    g.cand_bytes = (int32_t)(sizeof(float4) * cand.size());
    g.sph = sph.data();
    g.sph_sid = sid.data();
    g.D = D.data();
    g.T2 = T2.data();
    g.inv_T2 = inv_T2.data();
  }
};

// Validation and the grid's scalar parameters, shared by the host builder
// below and the device builder (nz_grid_gpu.hip): the cell counts and edges
// in g, the FP64 edges hd and the list margin M (safety margin plus slack for
// the tolerance band and the rounding of cell-relative coordinates).
// Returns 0, or -1 with a message in err.
inline int grid_params(const double* spheres, int64_t S, int32_t nstruct, const double* D,
                       const double* T2, const double L[3], double& cell_size, double& margin,
                       Grid& g, double hd[3], double& M, std::string& err) {
  if (S < 0 || (S > 0 && !spheres)) { err = "bad sphere array"; return -1; }
  if (S > (int64_t)INT32_MAX) { err = "more than 2^31 spheres"; return -1; }
  if (nstruct < 0 || !D || !T2) { err = "bad structure arrays"; return -1; }
  for (int a = 0; a < 3; ++a)
    if (!(L[a] > 0) || !std::isfinite(L[a])) { err = "box edges must be > 0"; return -1; }
  for (int k = 0; k <= nstruct; ++k)
    if (!(D[k] >= 0) || !(T2[k] > 0)) { err = "need D >= 0 and T2 > 0"; return -1; }
  for (int64_t i = 0; i < S; ++i) {
    const double sd = spheres[4 * S + i];
    if (!(spheres[3 * S + i] > 0) || sd < 0 || sd >= nstruct || sd != std::floor(sd)) {
      err = "sphere " + std::to_string(i) + " needs r > 0 and an integer sid in [0, nstructures)";
      return -1;
    }
  }
  if (cell_size <= 0) cell_size = 2.0;
  if (margin <= 0) margin = 0.8;
  std::memset(&g, 0, sizeof(g));
  g.nstruct = nstruct;
  g.nspheres = S;
  g.margin = (float)margin;
  for (int a = 0; a < 3; ++a) {
    g.ncell[a] = std::max(1, (int)std::lround(L[a] / cell_size));
    if (g.ncell[a] > 1000) { err = "more than 1000 cells per axis; raise cell_size"; return -1; }
    g.nloc[a] = g.ncell[a] + 2;
    g.h[a] = (float)(L[a] / g.ncell[a]);
    g.inv_h[a] = (float)(g.ncell[a] / L[a]);
    g.L[a] = (float)L[a];
    g.hd[a] = L[a] / g.ncell[a];
    g.Ld[a] = L[a];
    hd[a] = L[a] / g.ncell[a];
  }
  M = margin + 4.0 * (double)kTol + 1e-5;
  return 0;
}

// Returns 0, or -1 with a message in err.
inline int build_grid_host(const double* spheres, int64_t S, int32_t nstruct, const double* D,
                           const double* T2, const double L[3], double cell_size, double margin,
                           HostGrid& out, std::string& err) {
  Grid& g = out.g;
  double hd[3], M;
  if (grid_params(spheres, S, nstruct, D, T2, L, cell_size, margin, g, hd, M, err) != 0) return -1;
  std::memset(&out.info, 0, sizeof(out.info));
  const int64_t ncells = (int64_t)g.nloc[0] * g.nloc[1] * g.nloc[2];
  const double* cx = spheres;
  const double* cy = spheres + S;
  const double* cz = spheres + 2 * S;
  const double* r = spheres + 3 * S;
  const double* sd = spheres + 4 * S;
  auto cell_lo = [&](int a, int l) { return -L[a] / 2 + (l - 1) * hd[a]; };
  auto axis_range = [&](int64_t i, int a, int& lo, int& hi) {
    const double c = (a == 0 ? cx[i] : a == 1 ? cy[i] : cz[i]);
    lo = (int)std::floor((c - r[i] - M + L[a] / 2) / hd[a]) - 1;
    hi = (int)std::floor((c + r[i] + M + L[a] / 2) / hd[a]) + 2;
    lo = std::max(lo, 0);
    hi = std::min(hi, g.nloc[a] - 1);
    while (lo <= hi && cell_lo(a, lo) + hd[a] + M < c - r[i]) ++lo;  // exact AABB test
    while (hi >= lo && cell_lo(a, hi) - M > c + r[i]) --hi;
  };
  std::vector<int64_t> cnt(ncells + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int64_t> fill;
    if (pass == 1) {
      for (int64_t c = 0; c < ncells; ++c) cnt[c + 1] += cnt[c];
      if (cnt[ncells] > kMaxCandEntries) {
        err = "candidate lists exceed 2^26 entries (2 GiB of records); lower cell_size or margin";
        return -1;
      }
      // padding records after the last list: the scan loads ahead, up to
      // kCandPad - 1 records past the end of a list
      out.cand.assign(2 * ((size_t)cnt[ncells] + kCandPad), make_float4(0, 0, 0, 0));
      fill.assign(cnt.begin(), cnt.end() - 1);
    }
    for (int64_t i = 0; i < S; ++i) {
      int lo[3], hi[3];
      bool any = true;
      for (int a = 0; a < 3; ++a) {
        axis_range(i, a, lo[a], hi[a]);
        any = any && lo[a] <= hi[a];
      }
      if (!any) continue;
      const double ri = r[i];
      const int32_t idx = (int32_t)i;
      float idx_bits;
      std::memcpy(&idx_bits, &idx, sizeof(idx_bits));
      const float rout2 = (float)((ri + (double)kTol) * (ri + (double)kTol));
      const float4 bv = make_float4((float)(ri * ri),
                                    (float)((ri - (double)kTol) * (ri - (double)kTol)),
                                    (float)sd[i], idx_bits);
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) {
            const size_t c = ((size_t)z * g.nloc[1] + y) * g.nloc[0] + x;
            if (pass == 0) {
              cnt[c + 1]++;
            } else {
              const int64_t k = fill[c]++;
              out.cand[2 * k] = make_float4((float)(cx[i] - cell_lo(0, x)),
                                            (float)(cy[i] - cell_lo(1, y)),
                                            (float)(cz[i] - cell_lo(2, z)), rout2);
              out.cand[2 * k + 1] = bv;
            }
          }
    }
  }
  out.start.resize(ncells + 1);
  for (int64_t c = 0; c <= ncells; ++c) out.start[c] = (int32_t)cnt[c];
  {  // sort ranks of the in-box cells: R^3-cell regions, then list length
    const int R = NZ_RANK_REGION;
    const int nx = g.ncell[0], ny = g.ncell[1], nz = g.ncell[2];
    const int rx = (nx + R - 1) / R, ry = (ny + R - 1) / R;
    const int64_t nbox = (int64_t)nx * ny * nz;
    std::vector<int64_t> order(nbox);
    std::vector<int64_t> region(nbox), len(nbox);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          const int64_t c = ((int64_t)z * ny + y) * nx + x;
          const size_t lc = ((size_t)(z + 1) * g.nloc[1] + (y + 1)) * g.nloc[0] + (x + 1);
          region[c] = ((int64_t)(z / R) * ry + y / R) * rx + x / R;
          len[c] = cnt[lc + 1] - cnt[lc];
          order[c] = c;
        }
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      if (region[a] != region[b]) return region[a] < region[b];
      if (len[a] != len[b]) return len[a] < len[b];
      return a < b;
    });
    out.rank.assign(nbox, 0);
    for (int64_t k = 0; k < nbox; ++k) out.rank[order[k]] = (int32_t)k;
  }
  out.sph.resize(S);
  out.sid.resize(S);
  for (int64_t i = 0; i < S; ++i) {
    out.sph[i] = make_double4(cx[i], cy[i], cz[i], r[i]);
    out.sid[i] = (int32_t)sd[i];
  }
  out.D.assign(D, D + nstruct + 1);
  out.T2.assign(T2, T2 + nstruct + 1);
  out.inv_T2.resize(nstruct + 1);
  for (int32_t k = 0; k <= nstruct; ++k) out.inv_T2[k] = (float)(1.0 / T2[k]);

  nz_substrate_info_t& inf = out.info;
  inf.nspheres = S;
  inf.nstructures = nstruct;
  for (int a = 0; a < 3; ++a) {
    inf.ncell[a] = g.ncell[a];
    inf.cell[a] = hd[a];
  }
  inf.margin = margin;
  inf.ncand_total = cnt[ncells];
  inf.ncand_records = cnt[ncells] + kCandPad;
  int32_t mx = 0;
  double box_sum = 0;
  for (int z = 0; z < g.nloc[2]; ++z)
    for (int y = 0; y < g.nloc[1]; ++y)
      for (int x = 0; x < g.nloc[0]; ++x) {
        const size_t c = ((size_t)z * g.nloc[1] + y) * g.nloc[0] + x;
        const int32_t n = (int32_t)(cnt[c + 1] - cnt[c]);
        mx = std::max(mx, n);
        if (x >= 1 && x <= g.ncell[0] && y >= 1 && y <= g.ncell[1] && z >= 1 && z <= g.ncell[2])
          box_sum += n;
      }
  inf.ncand_max = mx;
  inf.ncand_mean_box = box_sum / ((double)g.ncell[0] * g.ncell[1] * g.ncell[2]);
  return 0;
}

}  // namespace nz
