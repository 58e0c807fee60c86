// Substrate upload and cell-grid construction (host side of libnozomi_hip.so).
//
// Replaces DiffSim3d.set_segments (diffsim3d.py:107-174).  The reference keeps,
// per segment, every sphere whose AABB is within 1 um of the segment and pads
// all lists to the longest one; the walker then uses the list of the pre-step
// segment for the whole step (sim3d_kernel.cu:367-369), so a jump longer than
// the margin can miss a sphere.  Here the grid tiles the box exactly, carries
// one halo cell per face for the image spheres beyond it, stores each list in
// CSR form with cell-relative single-precision coordinates, and the walker
// checks every sub-step's segment against the list's safety box (see
// nz_walk.hip), so the candidate set is always complete.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nz_internal.h"

namespace nz {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return (int)e;
}

template <class T>
static int upload(Substrate* s, const std::vector<T>& v, const T** out) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
  NZ_HIP(hipMalloc(&p, bytes));
  s->allocations.push_back(p);
  if (!v.empty()) NZ_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  s->info.device_bytes += (int64_t)bytes;
  *out = (const T*)p;
  return 0;
}

static void release(Substrate* s) {
  for (void* p : s->allocations) (void)hipFree(p);
  s->allocations.clear();
}

}  // namespace nz

using namespace nz;

extern "C" {

int nz_abi_version(void) { return NZ_ABI_VERSION; }

const char* nz_last_error_string(void) { return g_err.c_str(); }

int nz_substrate_create(const double* spheres, int64_t S, int32_t nstruct, const float* D,
                        const float* T2, double Lx, double Ly, double Lz, double cell_size,
                        double margin, nz_substrate_t* out) {
  if (!out) return fail(-1, "nz_substrate_create: out is NULL");
  *out = nullptr;
  if (S < 0 || (S > 0 && !spheres)) return fail(-1, "nz_substrate_create: bad sphere array");
  if (nstruct < 0 || !D || !T2) return fail(-1, "nz_substrate_create: bad structure arrays");
  const double L[3] = {Lx, Ly, Lz};
  for (int a = 0; a < 3; ++a)
    if (!(L[a] > 0) || !std::isfinite(L[a])) return fail(-1, "nz_substrate_create: box edges must be > 0");
  for (int k = 0; k <= nstruct; ++k)
    if (!(D[k] >= 0) || !(T2[k] > 0)) return fail(-1, "nz_substrate_create: need D >= 0 and T2 > 0");
  for (int64_t i = 0; i < S; ++i) {
    const double sid = spheres[4 * S + i];
    if (!(spheres[3 * S + i] > 0) || sid < 0 || sid >= nstruct || sid != std::floor(sid))
      return fail(-1, "nz_substrate_create: sphere " + std::to_string(i) +
                          " needs r > 0 and an integer sid in [0, nstructures)");
  }
  if (cell_size <= 0) cell_size = 2.0;
  if (margin <= 0) margin = 1.0;

  Substrate* s = new Substrate();
  std::memset(&s->info, 0, sizeof(s->info));
  (void)hipGetDevice(&s->device);
  Grid& g = s->g;
  g.nstruct = nstruct;
  g.nspheres = S;
  g.margin = (float)margin;
  int ncells = 1;
  for (int a = 0; a < 3; ++a) {
    g.ncell[a] = std::max(1, (int)std::lround(L[a] / cell_size));
    if (g.ncell[a] > 1000) {
      delete s;
      return fail(-1, "nz_substrate_create: more than 1000 cells per axis; raise cell_size");
    }
    g.nloc[a] = g.ncell[a] + 2;
    g.h[a] = (float)(L[a] / g.ncell[a]);
    g.inv_h[a] = (float)(g.ncell[a] / L[a]);
    g.L[a] = (float)L[a];
    ncells *= g.nloc[a];
  }
  const double hd[3] = {L[0] / g.ncell[0], L[1] / g.ncell[1], L[2] / g.ncell[2]};
  // list margin: the safety margin plus slack for the tolerance band and for
  // rounding of the cell-relative coordinates
  const double M = margin + 4.0 * (double)kTol + 1e-5;

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

  std::vector<int32_t> start(ncells + 1, 0);
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
        for (int x = lo[0]; x <= hi[0]; ++x) start[((size_t)z * g.nloc[1] + y) * g.nloc[0] + x + 1]++;
  }
  for (int c = 0; c < ncells; ++c) start[c + 1] += start[c];
  const int64_t total = start[ncells];
  if (total > (int64_t)INT32_MAX) {
    delete s;
    return fail(-1, "nz_substrate_create: candidate lists exceed 2^31 entries; lower cell_size/margin");
  }
  std::vector<float4> ca((size_t)total), cb((size_t)total);
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int64_t i = 0; i < S; ++i) {
    int lo[3], hi[3];
    bool any = true;
    for (int a = 0; a < 3; ++a) {
      axis_range(i, a, lo[a], hi[a]);
      any = any && lo[a] <= hi[a];
    }
    if (!any) continue;
    const double ri = r[i];
    const float4 bv = make_float4((float)(ri * ri), (float)((ri - (double)kTol) * (ri - (double)kTol)),
                                  (float)((ri + (double)kTol) * (ri + (double)kTol)), (float)ri);
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          const size_t c = ((size_t)z * g.nloc[1] + y) * g.nloc[0] + x;
          const int32_t k = fill[c]++;
          ca[k] = make_float4((float)(cx[i] - cell_lo(0, x)), (float)(cy[i] - cell_lo(1, y)),
                              (float)(cz[i] - cell_lo(2, z)), (float)sd[i]);
          cb[k] = bv;
        }
  }
  std::vector<double4> sph((size_t)S);
  std::vector<int32_t> sid((size_t)S);
  for (int64_t i = 0; i < S; ++i) {
    sph[i] = make_double4(cx[i], cy[i], cz[i], r[i]);
    sid[i] = (int32_t)sd[i];
  }
  std::vector<float> Dv(D, D + nstruct + 1), Tv(T2, T2 + nstruct + 1);

  int rc = 0;
  if ((rc = upload(s, start, &g.cell_start)) || (rc = upload(s, ca, &g.cand_a)) ||
      (rc = upload(s, cb, &g.cand_b)) || (rc = upload(s, sph, &g.sph)) ||
      (rc = upload(s, sid, &g.sph_sid)) || (rc = upload(s, Dv, &g.D)) ||
      (rc = upload(s, Tv, &g.T2))) {
    release(s);
    delete s;
    return rc;
  }

  nz_substrate_info_t& inf = s->info;
  inf.nspheres = S;
  inf.nstructures = nstruct;
  for (int a = 0; a < 3; ++a) {
    inf.ncell[a] = g.ncell[a];
    inf.cell[a] = hd[a];
  }
  inf.margin = margin;
  inf.ncand_total = total;
  int32_t mx = 0;
  double box_sum = 0;
  for (int z = 0; z < g.nloc[2]; ++z)
    for (int y = 0; y < g.nloc[1]; ++y)
      for (int x = 0; x < g.nloc[0]; ++x) {
        const size_t c = ((size_t)z * g.nloc[1] + y) * g.nloc[0] + x;
        const int32_t n = start[c + 1] - start[c];
        mx = std::max(mx, n);
        if (x >= 1 && x <= g.ncell[0] && y >= 1 && y <= g.ncell[1] && z >= 1 && z <= g.ncell[2]) box_sum += n;
      }
  inf.ncand_max = mx;
  inf.ncand_mean_box = box_sum / ((double)g.ncell[0] * g.ncell[1] * g.ncell[2]);
  *out = (nz_substrate_t)s;
  return 0;
}

int nz_substrate_destroy(nz_substrate_t sub) {
  if (!sub) return 0;
  Substrate* s = (Substrate*)sub;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(s->device);
  release(s);
  (void)hipSetDevice(dev);
  delete s;
  return 0;
}

int nz_substrate_info(nz_substrate_t sub, nz_substrate_info_t* info) {
  if (!sub || !info) return fail(-1, "nz_substrate_info: NULL argument");
  *info = ((Substrate*)sub)->info;
  return 0;
}

}  // extern "C"
