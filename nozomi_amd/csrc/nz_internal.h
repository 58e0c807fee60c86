// Internal types shared by the HIP kernels and the host side of
// libnozomi_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/nozomi_hip.h"

namespace nz {

// Tolerance band and "no hit" value of the reference walker
// (sim3d_kernel.cu:34-35) and its sub-step cap (:373).
constexpr float kTol = 1e-4f;
constexpr float kLarge = 1e6f;
constexpr int kMaxIter = 100;
// Longest candidate list the FP32 scan takes (it keys a list position in 8
// bits); a sub-step on a cell with a longer list scans it in FP64.
constexpr int kMaxListLen = 255;
// Rejection-seeding attempts per spin before it counts as unplaceable: a
// target filling 0.2% of the box (a 0.5 um cylinder in a 20 um box) needs
// ~500 on average, and a spin's attempts are keyed (k, 0, spin), so a larger
// bound changes no spin that an earlier one placed.
constexpr int32_t kSeedAttempts = 1 << 20;
// Zero records after the last candidate list (the scans load up to 15
// records past a list's end).
constexpr int kCandPad = 16;
// padding entries after the last reach list (the first pass loads up to 15
// entries past a list)
constexpr int kReachPad = 16;
// Most candidate entries of a substrate: the record array (32 bytes an entry,
// padding included) stays below 2 GiB, the range of the scan's buffer loads
// and of the int32 byte size it is described by.
constexpr int64_t kMaxCandEntries = ((int64_t)1 << 26) - kCandPad - 1;
static_assert((kMaxCandEntries + kCandPad) * 32 < ((int64_t)1 << 31),
              "record array size must fit in int32");

// Device view of a substrate.  The box [-L/2, L/2] per axis is tiled by
// ncell[a] cells of edge h[a]; one halo cell on each side (local index 0 and
// ncell+1) holds the explicit image spheres beyond the faces.  A cell's list
// holds every sphere whose AABB comes within `margin` (+ tol slack) of the
// cell box; a sub-step whose segment stays inside the cell box grown by
// `margin` therefore sees every sphere it can touch.
struct Grid {
  int32_t ncell[3];    // cells inside the box per axis
  int32_t nloc[3];     // ncell + 2 (with halo)
  float h[3];          // cell edge
  float inv_h[3];
  float L[3];          // box edge
  double hd[3];        // cell edge (FP64)
  double Ld[3];        // box edge (FP64)
  float margin;        // safety margin of the lists
  // the safety box [-margin, h + margin] per axis shrunk by 1e-5 (h + margin):
  // the FP32 test of a segment against it errs only towards "outside"
  float safe_lo;
  float safe_hi[3];
  // The reach test of the scan's first pass, per list entry (R = r + tol,
  // c the centre relative to the cell corner, ell the segment length, p the
  // position relative to the cell corner): a segment that touches the
  // R-ball has |p - c| <= R + ell, hence for any lambda > 0
  //   |p - c|^2 <= (1 + lambda) R^2 + (1 + 1/lambda) ell^2,
  // evaluated as p.h + k <= sigma ((1 + 1/lambda) ell^2 - |p|^2) + eps from
  // an 8-byte entry of FP16 values (h_x, h_y, h_z, k): h = -2 sigma c~ (c~
  // the centre h encodes), k <= sigma (|c~|^2 - (1 + lambda) R'^2) with R'
  // = R + |c - c~|.  sigma is a power of two that keeps the values in FP16
  // range; eps bounds the FP32 round-off of the test.
  float reach_sigma;
  float reach_lam1;    // sigma (1 + 1/lambda)
  float reach_eps;
  double reach_lambda;
  // entries of cell c's list from reach[reach_start[c]] (an even index:
  // two entries per 16-byte load), padding entries at the end
  const uint2* reach;
  const int32_t* reach_start;  // [nloc^3]
  int32_t reach_bytes;         // size of reach in bytes (< 2^31)
  // per cell (cell_start[c], cell_start[c + 1], reach_start[c], 0): what a
  // sub-step needs of its cell in one 16-byte load
  const int4* cell_tab;        // [nloc^3]
  int32_t nstruct;     // number of structures; the box has id nstruct
  int64_t nspheres;
  // CSR candidate lists over the nloc^3 local cells; entry e is the 32-byte
  // record cand[2e] = a, cand[2e+1] = b with cell-relative coordinates:
  //   a = (cx - lo_x, cy - lo_y, cz - lo_z, (r+tol)^2)
  //   b = (r^2, (r-tol)^2, sid, sphere index as int bits)
  // A list longer than kMaxListLen is scanned in FP64.
  const int32_t* cell_start;  // [ncells + 1]
  // sort position of each in-box cell (index (z ncell_y + y) ncell_x + x):
  // cells grouped in 4x4x4 regions, ordered by list length within a region,
  // so a wavefront's spins tend to scan lists of similar length
  const int32_t* cell_rank;
  const float4* cand;         // [2 * (entries + kCandPad)]: padding records at the end
  int32_t cand_bytes;         // size of cand in bytes (< 2^31)
  // every sphere, absolute, FP64: the exact re-solves and the rare sub-step
  // that leaves the grid's lists
  const double4* sph;         // (cx, cy, cz, r)
  const int32_t* sph_sid;
  // per-structure parameters, box last
  const double* D;
  const double* T2;
  const float* inv_T2;  // (float)(1 / T2), the factor of the log T2 weight
};

struct Substrate {
  Grid g;
  Grid* g_dev = nullptr;  // a copy of g in device memory
  nz_substrate_info_t info;
  std::vector<void*> allocations;
  int device = 0;
};

// thread-local error message
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Sum over the n entries of mom [3, ld] (entries with valid[i] < 0 skipped)
// of cos / sin of scale[w] * dirs[:, w] . m, into out [2W] (FP64): per-block
// FP64 partials into partial [signal_partial_len(n, W)], then summed over the
// blocks in block order.  _pooled takes the partials from the stream's pool.
int64_t signal_partial_len(int64_t n, int32_t W);
int signal_reduce(int64_t n, int64_t ld, const int32_t* valid, int32_t W, const float* mom,
                  const float* dirs, const float* scale, double* partial, double* out,
                  hipStream_t stream);
int signal_reduce_pooled(int64_t n, int64_t ld, const int32_t* valid, int32_t W, const float* mom,
                         const float* dirs, const float* scale, double* out, hipStream_t stream);

// In-place exclusive scan of counts[0, m) (int32; the total must fit), with
// a work array of scan_partial_len(m) ints; launches on `stream`.
int64_t scan_partial_len(int64_t m);
int exclusive_scan_i32(int32_t* counts, int64_t m, int32_t* partial, hipStream_t stream);

// Build the cell grid of a substrate on the device (nz_grid_gpu.hip): the
// same lists, records and sort ranks as build_grid_host (nz_grid_build.h).
int build_grid_device(const double* spheres, int64_t S, int32_t nstruct, const double* D,
                      const double* T2, const double L[3], double cell_size, double margin,
                      Substrate* s);

}  // namespace nz

#define NZ_HIP(call)                                           \
  do {                                                         \
    hipError_t _e = (call);                                    \
    if (_e != hipSuccess) return ::nz::hip_fail(_e, #call);    \
  } while (0)
