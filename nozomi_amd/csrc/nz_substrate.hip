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

static void release(Substrate* s) {
  for (void* p : s->allocations) (void)hipFree(p);
  s->allocations.clear();
}

}  // namespace nz

using namespace nz;

extern "C" {

int nz_abi_version(void) { return NZ_ABI_VERSION; }

const char* nz_last_error_string(void) { return g_err.c_str(); }

int nz_substrate_create(const double* spheres, int64_t S, int32_t nstruct, const double* D,
                        const double* T2, double Lx, double Ly, double Lz, double cell_size,
                        double margin, nz_substrate_t* out) {
  if (!out) return fail(-1, "nz_substrate_create: out is NULL");
  *out = nullptr;
  const double L[3] = {Lx, Ly, Lz};
  Substrate* s = new Substrate();
  (void)hipGetDevice(&s->device);
  if (int rc = build_grid_device(spheres, S, nstruct, D, T2, L, cell_size, margin, s)) {
    release(s);
    delete s;
    return rc;
  }
  *out = (nz_substrate_t)s;
  return 0;
}

int nz_substrate_export_grid(nz_substrate_t sub, int32_t* cell_start, int32_t* cell_rank,
                             float* cand) {
  if (!sub) return fail(-1, "nz_substrate_export_grid: NULL substrate");
  const Substrate* s = (const Substrate*)sub;
  const Grid& g = s->g;
  const size_t ncells = (size_t)g.nloc[0] * g.nloc[1] * g.nloc[2];
  const size_t nbox = (size_t)g.ncell[0] * g.ncell[1] * g.ncell[2];
  const size_t nrec = 2 * (size_t)s->info.ncand_records;
  if (cell_start)
    NZ_HIP(hipMemcpy(cell_start, g.cell_start, sizeof(int32_t) * (ncells + 1),
                     hipMemcpyDeviceToHost));
  if (cell_rank)
    NZ_HIP(hipMemcpy(cell_rank, g.cell_rank, sizeof(int32_t) * nbox, hipMemcpyDeviceToHost));
  if (cand) NZ_HIP(hipMemcpy(cand, g.cand, sizeof(float4) * nrec, hipMemcpyDeviceToHost));
  return 0;
}

int nz_substrate_export_reach(nz_substrate_t sub, int64_t* nentries, int32_t* reach_start,
                              uint32_t* reach) {
  if (!sub) return fail(-1, "nz_substrate_export_reach: NULL substrate");
  const Substrate* s = (const Substrate*)sub;
  const Grid& g = s->g;
  const size_t ncells = (size_t)g.nloc[0] * g.nloc[1] * g.nloc[2];
  const size_t n = (size_t)g.reach_bytes / sizeof(uint2);
  if (nentries) *nentries = (int64_t)n;
  if (reach_start)
    NZ_HIP(hipMemcpy(reach_start, g.reach_start, sizeof(int32_t) * ncells, hipMemcpyDeviceToHost));
  if (reach) NZ_HIP(hipMemcpy(reach, g.reach, sizeof(uint2) * n, hipMemcpyDeviceToHost));
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
