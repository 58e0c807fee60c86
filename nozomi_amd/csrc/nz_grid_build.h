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
  std::vector<uint2> reach;
  std::vector<int32_t> reach_start;
  std::vector<int4> cell_tab;
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
    g.reach = reach.data();
    g.reach_start = reach_start.data();
    g.reach_bytes = (int32_t)(sizeof(uint2) * reach.size());
    g.cell_tab = cell_tab.data();
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
  {
    double hmax = 0.0;
    for (int a = 0; a < 3; ++a) hmax = std::max(hmax, hd[a]);
    const double shrink = 1e-5 * (hmax + margin);
    g.safe_lo = (float)(-margin + shrink);
    for (int a = 0; a < 3; ++a) g.safe_hi[a] = (float)(hd[a] + margin - shrink);
  }
  M = margin + 4.0 * (double)kTol + 1e-5;
  // reach test parameters.  lambda = ell / R makes the bound exact; a
  // sub-step's segment is typically 0.4 of the list margin (the margin is
  // sized for the longer jumps), R the median radius + tol.
  double rmax = 0.0;
  std::vector<double> radii(spheres + 3 * S, spheres + 4 * S);
  for (double r : radii) rmax = std::max(rmax, r);
  double rmed = 1.0;
  if (S > 0) {
    std::nth_element(radii.begin(), radii.begin() + S / 2, radii.end());
    rmed = radii[S / 2];
  }
  const double lambda = std::min(4.0, std::max(0.05, 0.4 * margin / (rmed + (double)kTol)));
  // |c_a| <= h_a + M + r_max for a list entry (plus the FP16 rounding),
  // |p_a| <= h_a + M in the safety box, ell^2 below 4 p2
  const double R = rmax + (double)kTol;
  double c2 = 0.0, p2 = 0.0, cmax = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double ca = hd[a] + M + rmax;
    c2 += ca * ca;
    p2 += (hd[a] + M) * (hd[a] + M);
    cmax = std::max(cmax, ca);
  }
  const double kmax = 1.01 * (c2 + (1.0 + lambda) * (R + 0.01 * cmax) * (R + 0.01 * cmax));
  const double vmax = std::max(2.0 * cmax, kmax);
  int e2 = 0;
  (void)std::frexp(16384.0 / vmax, &e2);  // sigma = 2^(e2 - 1) <= 16384 / vmax
  const double sigma = std::ldexp(1.0, e2 - 1);
  g.reach_lambda = lambda;
  g.reach_sigma = (float)sigma;
  g.reach_lam1 = (float)(sigma * (1.0 + 1.0 / lambda));
  // the FP32 test's terms are below sigma B; six roundings stay below 4e-7 sigma B
  const double B = 2.0 * kmax + 2.0 * p2 + 4.0 * (1.0 + 1.0 / lambda) * p2;
  g.reach_eps = (float)(1e-6 * sigma * B);
  return 0;
}

// binary16 helpers for the reach entries: exact integer and power-of-two
// steps only (frexp, ldexp, rint), so the host and the device builder agree
// bit for bit.  Finite |x| < 65504.
__host__ __device__ inline double half_value(uint16_t h) {
  const int e = (h >> 10) & 31;
  const double m = (double)(h & 1023);
  const double v = e ? std::ldexp(1024.0 + m, e - 25) : std::ldexp(m, -24);
  return (h & 0x8000) ? -v : v;
}

// nearest binary16 (ties to even)
__host__ __device__ inline uint16_t half_rn(double x) {
  const uint16_t sign = std::signbit(x) ? 0x8000 : 0;
  const double a = std::fabs(x);
  if (a == 0.0) return sign;
  int e = 0;
  (void)std::frexp(a, &e);            // a in [2^(e-1), 2^e)
  int eh = std::max(e - 1, -14);      // binade of the result (subnormals: -14)
  double m = std::rint(std::ldexp(a, 10 - eh));  // a / ulp, rounded
  if (m >= 2048.0) {                  // rounded up into the next binade
    m *= 0.5;
    eh += 1;
  }
  const uint32_t mi = (uint32_t)m;
  if (mi < 1024u) return (uint16_t)(sign | mi);  // subnormal
  return (uint16_t)(sign | ((uint32_t)(eh + 15) << 10) | (mi - 1024u));
}

// largest binary16 <= x
__host__ __device__ inline uint16_t half_rd(double x) {
  uint16_t h = half_rn(x);
  if (half_value(h) > x) {
    if (h == 0) h = 0x8001;                 // +0 -> -2^-24
    else if (h & 0x8000) h = (uint16_t)(h + 1);  // more negative
    else h = (uint16_t)(h - 1);
  }
  return h;
}

// The reach entry of a sphere (centre c relative to the cell corner, radius
// r) for the grid's lambda and sigma: (h_x, h_y, h_z, k) as binary16 bits in
// two 32-bit words (h_x | h_y << 16, h_z | k << 16).
__host__ __device__ inline uint2 make_reach_entry(double cx, double cy, double cz, double r,
                                                  double lambda, double sigma) {
#pragma clang fp contract(off)
  const double c[3] = {cx, cy, cz};
  uint16_t h[3];
  double ct2 = 0.0, d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    h[a] = half_rn(-2.0 * sigma * c[a]);
    const double ct = half_value(h[a]) / (-2.0 * sigma);  // the centre h encodes (exact)
    ct2 += ct * ct;
    d2 += (c[a] - ct) * (c[a] - ct);
  }
  const double Rp = r + (double)kTol + std::sqrt(d2) * (1.0 + 1e-9);
  const double q = (1.0 + lambda) * Rp * Rp;
  // FP64 round-off of k before its downward rounding: below 1e-12 of the terms
  const double k = sigma * (ct2 - q) - 1e-12 * sigma * (ct2 + q);
  const uint16_t kh = half_rd(k);
  return make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)kh << 16));
}

// The two records of a list entry (centre relative to the cell's lower
// corner, radius, structure id, sphere index), shared by the host and the
// device builder.
__host__ __device__ inline void make_records(double cx, double cy, double cz, double r, double sid,
                                             int32_t idx, float4& a, float4& b) {
#pragma clang fp contract(off)
  const double tol = (double)kTol;
  float idx_bits;
  __builtin_memcpy(&idx_bits, &idx, sizeof(idx_bits));
  a = make_float4((float)cx, (float)cy, (float)cz, (float)((r + tol) * (r + tol)));
  b = make_float4((float)(r * r), (float)((r - tol) * (r - tol)), (float)sid, idx_bits);
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
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) {
            const size_t c = ((size_t)z * g.nloc[1] + y) * g.nloc[0] + x;
            if (pass == 0) {
              cnt[c + 1]++;
            } else {
              const int64_t k = fill[c]++;
              make_records(cx[i] - cell_lo(0, x), cy[i] - cell_lo(1, y), cz[i] - cell_lo(2, z),
                           r[i], sd[i], (int32_t)i, out.cand[2 * k], out.cand[2 * k + 1]);
            }
          }
    }
  }
  out.start.resize(ncells + 1);
  for (int64_t c = 0; c <= ncells; ++c) out.start[c] = (int32_t)cnt[c];
  {  // reach entries: each list from an even index, padding entries at the end
    out.reach_start.resize(ncells);
    int64_t rs = 0;
    for (int64_t c = 0; c < ncells; ++c) {
      out.reach_start[c] = (int32_t)rs;
      rs += (cnt[c + 1] - cnt[c] + 1) & ~(int64_t)1;
    }
    if ((rs + kReachPad) * 8 >= ((int64_t)1 << 31)) {
      err = "reach entries exceed 2 GiB; lower cell_size or margin";
      return -1;
    }
    out.reach.assign((size_t)(rs + kReachPad), make_uint2(0u, 0u));
    out.cell_tab.resize(ncells);
    for (int64_t c = 0; c < ncells; ++c)
      out.cell_tab[c] = make_int4((int32_t)cnt[c], (int32_t)cnt[c + 1], out.reach_start[c], 0);
    for (int z = 0; z < g.nloc[2]; ++z)
      for (int y = 0; y < g.nloc[1]; ++y)
        for (int x = 0; x < g.nloc[0]; ++x) {
          const int64_t c = ((int64_t)z * g.nloc[1] + y) * g.nloc[0] + x;
          for (int64_t e = cnt[c]; e < cnt[c + 1]; ++e) {
            int32_t i;
            std::memcpy(&i, &out.cand[2 * e + 1].w, sizeof(i));
            out.reach[out.reach_start[c] + (e - cnt[c])] =
                make_reach_entry(cx[i] - cell_lo(0, x), cy[i] - cell_lo(1, y),
                                 cz[i] - cell_lo(2, z), r[i], g.reach_lambda, g.reach_sigma);
          }
        }
  }
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
