// Per-spin walk logic, shared by the walk kernel (one spin per lane) and by
// the CPU build of the walker that the test suite checks against the oracle.
//
// Algorithm of the reference walker (sim3d_kernel.cu:340-522, with the phase
// additions of :524-717 and :719-943), restructured for gfx950:
//  * the candidate scan (the hot loop: every sphere of the cell list, every
//    sub-step) runs in FP32 on cell-relative coordinates.  It only has to
//    pick the decisive spheres: the containing sphere with the furthest exit
//    and the obstacle with the closest entry;
//  * a candidate whose FP32 inside test or discriminant falls within a guard
//    band of its decision threshold is re-solved in FP64, and the two winning
//    roots are always re-solved in FP64 with the reference's formulas, so the
//    branch taken and the contact point are those of a double-precision walk;
//  * the spin position, the jump vector and the jump bookkeeping are FP64 in
//    the reference's absolute frame, with its operation order and without
//    FP contraction: per sub-step, not per candidate, so they cost little.
//    Wall collisions amplify any rounding difference exponentially, so the
//    FP64 steps reproduce the double-precision restatement of the reference
//    operation for operation; only the FP32 candidate scan (which merely
//    picks the deciding spheres) works in cell-relative coordinates;
//  * the structure pass and the collision pass are fused into one scan that
//    assumes the structure id of the previous sub-step and re-runs only when
//    the id changed, which is exact with respect to the two-pass order;
//  * every sub-step checks that its segment stays inside the safety box of
//    its cell list (cell grown by the list margin) and re-bins or falls back
//    to all spheres when it does not, so no jump can miss a sphere;
//  * the T2 weight is kept as a log (one FMA per sub-step instead of an exp);
//  * phase is kept as the 3-vector moment sum_t g_t x_t: every direction of
//    the reference's multi-direction kernel is a projection of it, so the
//    per-step cost is independent of the number of directions.
#pragma once
#include <hip/hip_runtime.h>

#include "nz_internal.h"
#include "nz_rng.h"

namespace nz {

// the reference's comparison macros (sim3d_kernel.cu:7-8), NaN behaviour kept
NZ_HD inline float MINr(float a, float b) { return (a > b) ? b : a; }
NZ_HD inline float MAXr(float a, float b) { return (a < b) ? b : a; }
NZ_HD inline double MINd(double a, double b) { return (a > b) ? b : a; }
NZ_HD inline double MAXd(double a, double b) { return (a < b) ? b : a; }

// Guard bands of the FP32 decisions, relative to the magnitude of the terms
// that cancel: dd + r^2 for the inside test dd - r^2, and b^2 + 4a(dd + R^2)
// for the discriminant b^2 - 4a(dd - R^2) (c = dd - R^2 itself is pure
// round-off for a spin resting on a tolerance shell).  Single-precision
// round-off is ~1e-7 of those scales, two orders inside the bands.
constexpr float kGuardInside = 1e-5f;
constexpr float kGuardDisc = 2e-5f;

// Round-off band of an FP32 root as the scan keys it (the root itself to
// ~1e-5 relative, the key's 15 mantissa bits to 3e-5; a root near zero to
// ~1e-6 absolute): two roots closer than this cannot be ordered in FP32.
constexpr float kTieRel = 2e-4f;
constexpr float kTieAbs = 1e-5f;

// findStepSizeToSphereInsideTol (sim3d_kernel.cu:77-116) in FP64.
NZ_HD inline double step_inside_d(double dx, double dy, double dz, double r, double jx, double jy,
                                  double jz) {
#pragma clang fp contract(off)
  const double tol = (double)kTol;
  if ((dx * dx + dy * dy + dz * dz - r * r) > 0) return (double)kLarge;
  const double a = jx * jx + jy * jy + jz * jz;
  const double b = 2 * (jx * dx + jy * dy + jz * dz);
  const double c = dx * dx + dy * dy + dz * dz - (r - tol) * (r - tol);
  double sq = b * b - 4 * a * c;
  if (sq < 0) return 0;
  sq = sqrt(sq);
  double q = -0.5 * (b + sq);
  if (b < 0) q = -0.5 * (b - sq);
  return MAXd(MAXd(q / a, c / q), 0);
}

// findStepSizeToSphereOutsideTol (sim3d_kernel.cu:118-160) in FP64.
NZ_HD inline double step_outside_d(double dx, double dy, double dz, double r, double jx, double jy,
                                   double jz) {
#pragma clang fp contract(off)
  const double tol = (double)kTol;
  const double a = jx * jx + jy * jy + jz * jz;
  const double b = 2 * (jx * dx + jy * dy + jz * dz);
  const double c = dx * dx + dy * dy + dz * dz - (r + tol) * (r + tol);
  double sq = b * b - 4 * a * c;
  if (sq < 0) return (double)kLarge;
  sq = sqrt(sq);
  double q = -0.5 * (b + sq);
  if (b < 0) q = -0.5 * (b - sq);
  const double fpos = q / a, fneg = c / q;
  if ((dx * dx + dy * dy + dz * dz - r * r) < 0) return MAXd(fpos, fneg);
  if (fpos > 0 && fneg > 0) return MINd(fpos, fneg);
  if ((fpos < 0 && fneg < 0) || (b > 0)) return (double)kLarge;
  return 0;
}

// Both reference roots in one routine: findStepSizeToSphereInsideTol when
// `inside`, ...OutsideTol otherwise (sim3d_kernel.cu:77-160), same FP64
// arithmetic as step_inside_d / step_outside_d.
NZ_HD inline double root_d(bool inside, double dx, double dy, double dz, double r, double jx,
                           double jy, double jz) {
#pragma clang fp contract(off)
  const double tol = (double)kTol;
  const double d2r2 = dx * dx + dy * dy + dz * dz - r * r;
  if (inside && d2r2 > 0) return (double)kLarge;
  const double R = inside ? r - tol : r + tol;
  const double a = jx * jx + jy * jy + jz * jz;
  const double b = 2 * (jx * dx + jy * dy + jz * dz);
  const double c = dx * dx + dy * dy + dz * dz - R * R;
  double sq = b * b - 4 * a * c;
  if (sq < 0) return inside ? 0.0 : (double)kLarge;
  sq = sqrt(sq);
  double q = -0.5 * (b + sq);
  if (b < 0) q = -0.5 * (b - sq);
  const double fpos = q / a, fneg = c / q;
  if (inside) return MAXd(MAXd(fpos, fneg), 0);
  if (d2r2 < 0) return MAXd(fpos, fneg);
  if (fpos > 0 && fneg > 0) return MINd(fpos, fneg);
  if ((fpos < 0 && fneg < 0) || (b > 0)) return (double)kLarge;
  return 0;
}

struct ScanAcc {
  float sid_new;  // smallest id of a containing sphere (box id if none)
  float fin;      // furthest exit fraction among containing spheres
  float fout;     // closest entry fraction among lower-id spheres
  int kin;        // sphere index of fin (-1: none above 0)
  int kout;       // sphere index of fout (-1: none)
  // runners-up: the second furthest exit and the second closest entry (an
  // FP32 root cannot order two roots closer than its own round-off, so both
  // are re-solved in FP64 when they are that close)
  float fin2, fout2;
  int kin2, kout2;  // -1: none
  bool any_in;    // inside some relevant sphere
  bool exact;     // a candidate fell in a guard band: redo the scan in FP64
};

// Lane-invariant inputs of the FP32 candidate loop.
struct ScanIn {
  float p[3], j[3];
  float a4, inv_a, sid;  // 4 |j|^2, 1 / |j|^2, structure id assumed for the scan
  // right-hand side of the reach test (scan_list): with ell the sub-step's
  // segment length plus slack, sigma ((1 + 1/lambda) ell^2 - |p|^2) + eps
  float rhs;
};

// Accumulators of the FP32 candidate loop.  A decision and the list position
// of the candidate that made it share one 32-bit key: the root's float bits
// (non-negative, so they order as integers) with the low 8 mantissa bits
// replaced by the position + 1.  The furthest exit is the largest key, the
// closest entry the smallest; the roots keep 2^-15 relative precision, which
// is all the later FP64 re-solve needs to know which roots can decide.
// Candidates inside a guard band are left out and counted; the one position
// is kept for an FP64 visit after the loop.
struct ScanKeys {
  uint32_t kin;    // 0: no relevant containing sphere
  uint32_t kout;   // 0xffffffff: no obstacle
  uint32_t kin2;   // second largest exit key (0: none)
  uint32_t kout2;  // second smallest entry key (0xffffffff: none)
  float sid_new;   // smallest id of a containing sphere (box id if none)
  int32_t gcnt;    // candidates in a guard band
  uint32_t gpos;   // position + 1 of the last of them
};
constexpr uint32_t kKeyMask = 0x7fffff00u;  // list positions (< kMaxListLen) in the low 8 bits

NZ_HD inline uint32_t f2u(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, sizeof(u));
  return u;
}
NZ_HD inline float u2f(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, sizeof(f));
  return f;
}

// Unsigned max / min (one v_max_u32 / v_min_u32 each).
NZ_HD inline uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
NZ_HD inline uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

NZ_HD inline void acc_reset(const Grid& g, ScanAcc& acc) {
  acc.sid_new = (float)g.nstruct;
  acc.fin = 0.0f;
  acc.fout = kLarge;
  acc.kin = -1;
  acc.kout = -1;
  acc.fin2 = 0.0f;
  acc.fout2 = kLarge;
  acc.kin2 = -1;
  acc.kout2 = -1;
  acc.any_in = false;
  acc.exact = false;
}

// One candidate in FP64 (sphere index k; x the absolute position, j the
// jump): the exact structure and collision tests of the reference
// (isInsideSphere and the two root routines, sim3d_kernel.cu:72-160,380-455).
NZ_HD inline void visit_exact(const Grid& g, const double x[3], const double j[3], float sid, int k,
                              ScanAcc& acc) {
#pragma clang fp contract(off)
  const double4 s = g.sph[k];
  const float sidc = (float)g.sph_sid[k];
  const double dx = x[0] - s.x;
  const double dy = x[1] - s.y;
  const double dz = x[2] - s.z;
  const bool inside = (dx * dx + dy * dy + dz * dz - s.w * s.w) < 0;
  if (inside && sidc < acc.sid_new) acc.sid_new = sidc;
  if (sidc > sid) return;
  if (inside) {
    const float f = (float)step_inside_d(dx, dy, dz, s.w, j[0], j[1], j[2]);
    acc.any_in = true;
    if (acc.fin < f) {
      acc.fin2 = acc.fin;
      acc.kin2 = acc.kin;
      acc.fin = f;
      acc.kin = k;
    } else if (acc.fin2 < f) {
      acc.fin2 = f;
      acc.kin2 = k;
    }
  } else if (sidc < sid) {
    const float f = (float)step_outside_d(dx, dy, dz, s.w, j[0], j[1], j[2]);
    if (f < acc.fout) {
      acc.fout2 = acc.fout;
      acc.kout2 = acc.kout;
      acc.fout = f;
      acc.kout = k;
    } else if (f < acc.fout2) {
      acc.fout2 = f;
      acc.kout2 = k;
    }
  }
}

// One candidate sphere: the structure test (sim3d_kernel.cu:380-398) and the
// collision test (:414-455) fused, without branches.  A = (cx, cy, cz,
// (r+tol)^2) relative to the lane's cell corner, B = (r^2, (r-tol)^2, sid,
// sphere index bits).  One quadratic a f^2 + b f + c (c = dd - (r -/+ tol)^2) serves
// both reference roots (findStepSizeToSphereInsideTol / ...OutsideTol,
// :77-160), whose branch rules reduce to: with q = -(b - sqrt(disc))/2 for
// b <= 0 (q >= 0, far root q/a, near root c/q) and q = -(b + sqrt(disc))/2
// for b > 0,
//   inside : max(0, b <= 0 ? q/a : c/q), 0 on a miss;
//   outside: b > 0 or a miss -> no hit, else max(0, c/q).
NZ_HD inline void visit(const ScanIn& in, const float4 A, const float4 B, uint32_t pos1,
                        ScanKeys& k) {
  const float dx = in.p[0] - A.x, dy = in.p[1] - A.y, dz = in.p[2] - A.z;
  const float dd = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  const float din = dd - B.x;
  const float hb = fmaf(in.j[2], dz, fmaf(in.j[1], dy, in.j[0] * dx));
  const float b = hb + hb;
  const bool inside = din < 0.0f;                              // isInsideSphere < 0
  const float R2 = inside ? B.y : A.w;
  const float c = dd - R2;
  const float b2 = b * b;
  const float disc = fmaf(-in.a4, c, b2);
  const bool relevant = B.z <= in.sid;                         // :418-419
  // FP32 cannot decide this candidate: it is re-solved in FP64 after the loop
  // (bitwise bool algebra keeps the masks in scalar registers)
  const bool g_in = fabsf(din) < kGuardInside * (dd + B.x);
  const bool g_disc = fabsf(disc) < kGuardDisc * fmaf(in.a4, dd + R2, b2);
  const bool guard = g_in | (relevant & g_disc);
  k.gcnt += guard ? 1 : 0;
  k.gpos = guard ? pos1 : k.gpos;
  const bool take_sid = inside & !guard & (B.z < k.sid_new);  // :384-396
  k.sid_new = take_sid ? B.z : k.sid_new;
  const float sq = nz_sqrtf(fmaxf(disc, 0.0f));
  const bool away = b > 0.0f;
  const float q = -0.5f * (b + (away ? sq : -sq));
  const float near = c * nz_rcpf(q);
  const float root = (inside & !away) ? q * in.inv_a : near;
  const bool miss = disc < 0.0f;
  const float f_in = miss ? 0.0f : fmaxf(root, 0.0f);
  const float f_out = (miss | away) ? kLarge : fmaxf(near, 0.0f);
  const uint32_t key_in = (f2u(f_in) & kKeyMask) | pos1;
  const uint32_t key_out = (f2u(f_out) & kKeyMask) | pos1;
  const bool use_in = relevant & inside & !guard;                     // :428-440
  const bool use_out = !inside & (B.z < in.sid) & !guard;             // :441-454
  // best and runner-up by integer max / min (0 and 0xffffffff never win)
  const uint32_t ki = use_in ? key_in : 0u;
  const uint32_t ko = use_out ? key_out : 0xffffffffu;
  k.kin2 = umax(k.kin2, umin(k.kin, ki));
  k.kin = umax(k.kin, ki);
  k.kout2 = umin(k.kout2, umax(k.kout, ko));
  k.kout = umin(k.kout, ko);
}

NZ_HD inline int cell_id(const Grid& g, int lx, int ly, int lz) {
  return (lz * g.nloc[1] + ly) * g.nloc[0] + lx;
}

NZ_HD inline double cell_lo(const Grid& g, int a, int l) {
  return -0.5 * g.Ld[a] + (double)(l - 1) * g.hd[a];
}

// Segment p -> p + jr*j (the FP32 scan inputs) inside the lane cell's safety
// box, shrunk by far more than the FP32 round-off of p and p + jr j (below
// 1e-6 of the cell scale): a segment it accepts is inside the box.
NZ_HD inline bool in_safety_f(const Grid& g, const ScanIn& in, float jr) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float e = fmaf(jr, in.j[a], in.p[a]);
    ok = ok & (in.p[a] >= g.safe_lo) & (in.p[a] <= g.safe_hi[a]) & (e >= g.safe_lo) &
         (e <= g.safe_hi[a]);
  }
  return ok;
}

// Move the lane to the cell holding its position (in-step re-bin); rel, the
// position relative to the cell corner, follows.  Returns false when that
// cell is outside the grid's halo.
NZ_HD inline bool rebin(const Grid& g, double rel[3], int l[3]) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int k = (int)floor(rel[a] / g.hd[a]);
    if (k != 0) {
      const int nl = l[a] + k;
      if (nl < 0 || nl > g.ncell[a] + 1) {
        ok = false;
      } else {
        l[a] = nl;
        rel[a] -= (double)k * g.hd[a];
      }
    }
  }
  return ok;
}

NZ_HD inline void keys_reset(const Grid& g, ScanKeys& k) {
  k.kin = 0u;
  k.kout = 0xffffffffu;
  k.kin2 = 0u;
  k.kout2 = 0xffffffffu;
  k.sid_new = (float)g.nstruct;
  k.gcnt = 0;
  k.gpos = 0u;
}

// Sphere index of list entry e (the B record's w bits).
NZ_HD inline int cand_sphere(const Grid& g, int e) {
  return (int)f2u(g.cand[2 * (int64_t)e + 1].w);
}

// Loads of the records of one candidate list (position e = record cs + e).
// On the device these are buffer loads on the whole record array: a plain
// load carried from one loop trip to the next is merged by the compiler into
// a load at its use, which would undo the loads issued a trip ahead; buffer
// loads are kept where they are written.  Every offset stays inside the
// array (padding records included).
#if defined(__HIP_DEVICE_COMPILE__)
// Keeps the instruction scheduler from moving code across this point (the
// loads issued above stay ahead of the arithmetic below).
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }
// A buffer resource over [p, p + bytes).  Both are the same on every lane
// (one grid per launch); taking them from the first lane keeps the resource
// in scalar registers where the compiler cannot prove that itself (it would
// otherwise loop over the distinct resources of the wavefront at every load).
NZ_HD inline __amdgpu_buffer_rsrc_t uniform_rsrc(const void* p, int32_t bytes) {
  const uint64_t v = (uint64_t)p;
  // readfirstlane returns int: widen through uint32_t (a sign extension of
  // the low word would overwrite the high word when bit 31 is set)
  const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return __builtin_amdgcn_make_buffer_rsrc((void*)((hi << 32) | lo), (short)0,
                                           __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
typedef unsigned int nz_u32x4 __attribute__((ext_vector_type(4)));
// The buffer resources of the candidate records and the reach entries,
// built once per kernel where every lane is active (they stay in scalar
// registers; built inside the scan, their loads and first-lane reads would
// sit on every sub-step's critical path).
struct ListBufs {
  __amdgpu_buffer_rsrc_t cand, reach;
};
NZ_HD inline ListBufs list_bufs(const Grid& g) {
  return ListBufs{uniform_rsrc(g.cand, g.cand_bytes), uniform_rsrc(g.reach, g.reach_bytes)};
}
struct RecLoader {
  __amdgpu_buffer_rsrc_t r;
  int base;  // byte offset of the list's first record
  __device__ RecLoader(const ListBufs& b, int cs) : r(b.cand), base(cs * 32) {}
  __device__ float4 ld(int voff, int soff) const {
    const nz_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                       __uint_as_float(v.w));
  }
  // position e, the same on every lane (its offset goes in a scalar register)
  __device__ float4 a_u(int e) const { return ld(base, e * 32); }
  // position e of this lane
  __device__ float4 a(int e) const { return ld(base + e * 32, 0); }
  __device__ float4 b(int e) const { return ld(base + e * 32 + 16, 0); }
};
// Loads of the reach entries of one list (from entry rs, an even index).
struct ReachLoader {
  __amdgpu_buffer_rsrc_t r;
  int base;  // byte offset of the list's first entry (16-byte aligned)
  __device__ ReachLoader(const ListBufs& b, int rs) : r(b.reach), base(rs * 8) {}
  // entries e and e + 1 (e even, the same on every active lane: the loop
  // counter of the scan, taken from the first lane so the offset is scalar)
  __device__ nz_u32x4 ld2(int e) const {
    return __builtin_amdgcn_raw_buffer_load_b128(r, base, __builtin_amdgcn_readfirstlane(e * 8),
                                                 0);
  }
};
#else
NZ_HD inline void sched_fence() {}
struct ListBufs {
  const float4* cand;
  const uint2* reach;
};
NZ_HD inline ListBufs list_bufs(const Grid& g) { return ListBufs{g.cand, g.reach}; }
struct RecLoader {
  const float4* p;
  RecLoader(const ListBufs& b, int cs) : p(b.cand + 2 * (int64_t)cs) {}
  float4 a_u(int e) const { return p[2 * e]; }
  float4 a(int e) const { return p[2 * e]; }
  float4 b(int e) const { return p[2 * e + 1]; }
};
struct Words4 {
  uint32_t x, y, z, w;
};
struct ReachLoader {
  const uint2* p;
  ReachLoader(const ListBufs& b, int rs) : p(b.reach + rs) {}
  Words4 ld2(int e) const { return Words4{p[e].x, p[e].y, p[e + 1].x, p[e + 1].y}; }
};
#endif

// Every candidate of the list through the full visit, with the next record's
// loads issued before the current one is evaluated.  Loads past the list
// stay in bounds: the record array ends with padding records.
NZ_HD inline void scan_all(const float4* rec, int n, const ScanIn& in, ScanKeys& k) {
  float4 A0 = rec[0];
  float4 B0 = rec[1];
  int e = 0;
  for (; e + 2 <= n; e += 2) {  // two records per trip, each loaded one visit ahead
    const float4 A1 = rec[2];
    const float4 B1 = rec[3];
    visit(in, A0, B0, (uint32_t)(e + 1), k);
    A0 = rec[4];
    B0 = rec[5];
    visit(in, A1, B1, (uint32_t)(e + 2), k);
    rec += 4;
  }
  if (e < n) visit(in, A0, B0, (uint32_t)(e + 1), k);
}

// Can the sub-step's segment (start p, length at most ell) touch the
// candidate's R-ball (R = r + tol)?  Only if |p - c| <= R + ell; the reach
// entry (Grid::reach) turns the weaker |p - c|^2 <= (1 + lambda) R^2 +
// (1 + 1/lambda) ell^2 into three FMAs on its FP16 values, with every
// rounding on the side of passing.  A sphere that fails neither contains the
// spin nor can be entered before jr: it cannot change the exit key, the entry
// key within jr or the structure id.  w0 = h_x | h_y << 16, w1 = h_z | k << 16.
#if defined(__HIP_DEVICE_COMPILE__)
typedef _Float16 nz_h2 __attribute__((ext_vector_type(2)));
NZ_HD inline bool reach_entry(const ScanIn& in, uint32_t w0, uint32_t w1) {
  const nz_h2 a = __builtin_bit_cast(nz_h2, w0);
  const nz_h2 b = __builtin_bit_cast(nz_h2, w1);
  // FP16 operands of FP32 FMAs: v_fma_mix_f32, no conversions
  float t = fmaf(in.p[0], (float)a.x, (float)b.y);
  t = fmaf(in.p[1], (float)a.y, t);
  t = fmaf(in.p[2], (float)b.x, t);
  return t <= in.rhs;
}
#else
inline float half_to_float(uint32_t h) {  // exact
  const int e = (int)((h >> 10) & 31u);
  const float m = (float)(h & 1023u);
  const float v = e ? std::ldexp(1024.0f + m, e - 25) : std::ldexp(m, -24);
  return (h & 0x8000u) ? -v : v;
}
inline bool reach_entry(const ScanIn& in, uint32_t w0, uint32_t w1) {
  float t = fmaf(in.p[0], half_to_float(w0 & 0xffffu), half_to_float(w1 >> 16));
  t = fmaf(in.p[1], half_to_float(w0 >> 16), t);
  t = fmaf(in.p[2], half_to_float(w1 & 0xffffu), t);
  return t <= in.rhs;
}
#endif

// Reach bits (bit i: entry i) of eight entries, two per 16-byte word group.
template <class W>
NZ_HD inline uint32_t reach8(const ScanIn& in, const W& a, const W& b, const W& c, const W& d) {
  return (uint32_t)reach_entry(in, a.x, a.y) | ((uint32_t)reach_entry(in, a.z, a.w) << 1) |
         ((uint32_t)reach_entry(in, b.x, b.y) << 2) | ((uint32_t)reach_entry(in, b.z, b.w) << 3) |
         ((uint32_t)reach_entry(in, c.x, c.y) << 4) | ((uint32_t)reach_entry(in, c.z, c.w) << 5) |
         ((uint32_t)reach_entry(in, d.x, d.y) << 6) | ((uint32_t)reach_entry(in, d.z, d.w) << 7);
}

// Lowest set position of the mask, cleared; `none` when it is empty.
NZ_HD inline int pop_flag(uint64_t& m, int none) {
  if (m == 0u) return none;
  const int q = __builtin_ctzll(m);
  m &= m - 1u;
  return q;
}

// The candidate scan over the list [cs, ce) (reach entries from rs).  A
// reach pass over every candidate (its 8-byte reach entry) marks the spheres
// the segment can touch as bits of a 64-bit mask (the loop position is the
// same on every lane); only those get the full visit, a few per list against
// ~20 candidates.  Against visiting every candidate the exit key and the
// structure id are equal; the entry key can only differ by entries beyond jr
// (the skipped obstacles'), which cannot decide the sub-step, and the guard
// count by out-of-reach spheres, whose FP64 re-solve could not change it
// either.  Lists longer than 64 take the full visit for every candidate.
NZ_HD inline void scan_list(const Grid& g, const ListBufs& lb, int cs, int ce, int rs,
                            const ScanIn& in, ScanKeys& k) {
  const int n = ce - cs;
  if (n > 64) {
    scan_all(g.cand + 2 * (int64_t)cs, n, in, k);
    return;
  }
  const RecLoader rl(lb, cs);
  const ReachLoader ql(lb, rs);
  // eight entries (four 16-byte loads) per group, two groups alternating:
  // one group's loads are in flight while the other is tested, and the
  // register sets swap roles, so nothing is copied.  Positions past the list
  // (loads stay below n + 16) are tested too and masked off after the loop.
  uint64_t m = 0u;
  auto A0 = ql.ld2(0), A1 = ql.ld2(2), A2 = ql.ld2(4), A3 = ql.ld2(6);
  int e = 0;
  while (true) {
    const auto N0 = ql.ld2(e + 8), N1 = ql.ld2(e + 10), N2 = ql.ld2(e + 12), N3 = ql.ld2(e + 14);
    sched_fence();
    m |= (uint64_t)reach8(in, A0, A1, A2, A3) << e;
    e += 8;
    if (e >= n) break;
    A0 = ql.ld2(e + 8);
    A1 = ql.ld2(e + 10);
    A2 = ql.ld2(e + 12);
    A3 = ql.ld2(e + 14);
    sched_fence();
    m |= (uint64_t)reach8(in, N0, N1, N2, N3) << e;
    e += 8;
    if (e >= n) break;
  }
  m &= n >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1u);
  // the marked records, each loaded while the previous one is visited, into
  // two register sets that alternate (nothing is copied); the position n
  // (none left) loads the record after the list, not visited
  int q = pop_flag(m, n);
  float4 A = rl.a(q);
  float4 B = rl.b(q);
  while (q < n) {
    const int q2 = pop_flag(m, n);
    const float4 A2 = rl.a(q2);
    const float4 B2 = rl.b(q2);
    visit(in, A, B, (uint32_t)(q + 1), k);
    if (q2 >= n) break;
    q = pop_flag(m, n);
    A = rl.a(q);
    B = rl.b(q);
    visit(in, A2, B2, (uint32_t)(q2 + 1), k);
  }
}

// The keys of a finished scan as the FP64 code's accumulator.
NZ_HD inline void keys_decode(const Grid& g, int cs, const ScanKeys& k, ScanAcc& acc) {
  acc.sid_new = k.sid_new;
  acc.any_in = k.kin != 0u;
  acc.fin = acc.any_in ? u2f(k.kin & kKeyMask) : 0.0f;
  acc.kin = acc.any_in ? cand_sphere(g, cs + (int)(k.kin & 0xffu) - 1) : -1;
  const bool any_out = k.kout != 0xffffffffu;
  acc.fout = any_out ? u2f(k.kout & kKeyMask) : kLarge;
  acc.kout = any_out ? cand_sphere(g, cs + (int)(k.kout & 0xffu) - 1) : -1;
  // the runners-up (sphere indices as for the best)
  acc.fin2 = k.kin2 != 0u ? u2f(k.kin2 & kKeyMask) : -1.0f;
  acc.kin2 = k.kin2 != 0u ? cand_sphere(g, cs + (int)(k.kin2 & 0xffu) - 1) : -1;
  acc.fout2 = k.kout2 != 0xffffffffu ? u2f(k.kout2 & kKeyMask) : kLarge;
  acc.kout2 = k.kout2 != 0xffffffffu ? cand_sphere(g, cs + (int)(k.kout2 & 0xffu) - 1) : -1;
  acc.exact = false;
}

// The scan in FP64 (exact re-solve of a sub-step): over the lane's cell list
// [cs, ce) (mode 0), over the lists of the 27 cells around the lane's cell
// (mode 1: a segment that leaves the cell's safety box but not its
// neighbours'), or over every sphere (mode 2).
enum { kExactList = 0, kExactNeighbours = 1, kExactAll = 2 };

// The exact scan is inlined by default; building with
// -DNZ_EXACT_ATTR=__noinline__ moves it out of line (a call costs the kernel
// the callee's registers, so inlining is usually the better trade).
#ifndef NZ_EXACT_ATTR
#define NZ_EXACT_ATTR inline
#endif

// The geometry fields the exact scan reads, passed by value.
struct ExactGeom {
  const double4* sph;
  const int32_t* sph_sid;
  const int32_t* cell_start;
  const float4* cand;
  int64_t nspheres;
  int32_t nloc0, nloc1;
  int32_t nstruct;
};

NZ_HD NZ_EXACT_ATTR ScanAcc scan_exact_call(ExactGeom eg, int mode, int cs, int ce, double px,
                                           double py, double pz, double jx, double jy, double jz,
                                           int lx, int ly, int lz, float sid) {
  const double p[3] = {px, py, pz};
  const double j[3] = {jx, jy, jz};
  Grid g;  // only the fields visit_exact and acc_reset read
  g.sph = eg.sph;
  g.sph_sid = eg.sph_sid;
  g.nstruct = eg.nstruct;
  ScanAcc acc;
  acc_reset(g, acc);
  if (mode == kExactAll) {
    for (int64_t k = 0; k < eg.nspheres; ++k) visit_exact(g, p, j, sid, (int)k, acc);
    return acc;
  }
  const int r = (mode == kExactNeighbours) ? 1 : 0;
  for (int dz = -r; dz <= r; ++dz)
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx) {
        int b = cs, e = ce;
        if (mode == kExactNeighbours) {
          const int c = ((lz + dz) * eg.nloc1 + (ly + dy)) * eg.nloc0 + (lx + dx);
          b = eg.cell_start[c];
          e = eg.cell_start[c + 1];
        }
        for (int q = b; q < e; ++q) {
          int k;
          const float w = eg.cand[2 * (int64_t)q + 1].w;
          __builtin_memcpy(&k, &w, sizeof(k));
          visit_exact(g, p, j, sid, k, acc);
        }
      }
  return acc;
}

NZ_HD inline void scan_exact(const Grid& g, int mode, int cs, int ce, const double x[3],
                             const double j[3], const int l[3], float sid, ScanAcc& acc) {
  ExactGeom eg;
  eg.sph = g.sph;
  eg.sph_sid = g.sph_sid;
  eg.cell_start = g.cell_start;
  eg.cand = g.cand;
  eg.nspheres = g.nspheres;
  eg.nloc0 = g.nloc[0];
  eg.nloc1 = g.nloc[1];
  eg.nstruct = g.nstruct;
  acc = scan_exact_call(eg, mode, cs, ce, x[0], x[1], x[2], j[0], j[1], j[2], l[0], l[1], l[2],
                        sid);
}

// Segment inside the union of the safety boxes of the 27 cells around the
// lane's cell, all of which exist?
NZ_HD inline bool in_neighbourhood(const Grid& g, const double p[3], const double j[3], double jr,
                                   const int l[3]) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double e = p[a] + jr * j[a];
    const double lo = -g.hd[a] - (double)g.margin, hi = 2.0 * g.hd[a] + (double)g.margin;
    ok = ok && l[a] >= 1 && l[a] <= g.ncell[a] && p[a] >= lo && p[a] <= hi && e >= lo && e <= hi;
  }
  return ok;
}

// Resident per-spin state held in registers across the steps of a launch.
struct Lane {
  double x[3];   // absolute position, wrapped at step ends as the reference wraps
  int l[3];      // local index of the cell holding x (1..ncell inside the box,
                 // 0 / ncell+1 halo)
  int w[3];      // net +L shifts of the start position (the reference's spins0
                 // follows every wrap: spins0 = start + w L)
  float logsig;  // log T2 weight
  float m[3];    // phase moment
  float sid;     // structure id after the last sub-step
  int nsub;      // scatter sub-steps
  int nguard;    // sub-steps whose scan ran in FP64
};

// Position relative to the lower corner of the lane's cell (FP32 scan input
// and the safety-box test; the FP64 decisions use x itself).
NZ_HD inline void lane_rel(const Grid& g, const Lane& s, double rel[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) rel[a] = s.x[a] - cell_lo(g, a, s.l[a]);
}

// The cell holding coordinate x on axis a, clamped to the grid's halo.
NZ_HD inline int cell_of(const Grid& g, int a, double x) {
  int c = (int)floor((x + 0.5 * g.Ld[a]) / g.hd[a]) + 1;
  return c < 0 ? 0 : (c > g.ncell[a] + 1 ? g.ncell[a] + 1 : c);
}

// v as a value the compiler cannot see through: the redraw of the first
// jump after a structure change depends on nothing computed in the sub-step,
// and without this the compiler evaluates its normals speculatively on every
// sub-step (about 230 VALU) instead of in the rare branch that needs them.
NZ_HD inline uint32_t opaque_u32(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__("" : "+v"(v));
#endif
  return v;
}

NZ_HD inline void draw_jump(uint64_t seed, uint32_t spin, uint32_t step, uint32_t draw, double ds,
                            double j[3]) {
  float n0, n1, n2;
  normals3(seed, kStreamWalk, spin, step, draw, n0, n1, n2);
  j[0] = ds * (double)n0;
  j[1] = ds * (double)n1;
  j[2] = ds * (double)n2;
}

// Progress of one spin through its current time step.
struct StepState {
  double jr;    // remaining jump fraction (sim3d_kernel.cu:364)
  double ds;    // jump scale sqrt(2 D dt) of the step
  float ds_sid; // structure id that ds belongs to (-1: none yet; dt is fixed)
  double j[3];  // current jump vector
  int niter;    // sub-steps done in this step
  int cs, ce;   // candidate list of the lane's cell
  int rs;       // its first reach entry
  int pending;  // draw index of the jump still to draw before the next sub-step (-1: none)
};

// The FP32 scan's inputs for position rel (relative to the lane's cell
// corner), jump j, remaining fraction jr and assumed structure id sid.
NZ_HD inline void scan_in(const Grid& g, const double rel[3], const double j[3], double jr,
                          float sid, ScanIn& in) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    in.p[a] = (float)rel[a];
    in.j[a] = (float)j[a];
  }
  const float a = in.j[0] * in.j[0] + in.j[1] * in.j[1] + in.j[2] * in.j[2];
  in.a4 = 4.0f * a;
  // 1/a serves the FP32 roots only (their decisions carry guard and tie
  // bands far above the hardware reciprocal's error)
  in.inv_a = nz_rcpf(a);
  in.sid = sid;
  // segment length with slack (relative 1e-3, absolute 1e-4 um) for the
  // FP32 round-off of its length
  const float ell = nz_sqrtf(a) * (float)jr * 1.001f + 1e-4f;
  const float pp = fmaf(in.p[2], in.p[2], fmaf(in.p[1], in.p[1], in.p[0] * in.p[0]));
  in.rhs = fmaf(g.reach_lam1, ell * ell, g.reach_eps) - g.reach_sigma * pp;
}

// Start a time step.  The first jump (draw 0) is drawn with the structure
// id of the previous sub-step; substep() redraws it if the structure pass
// disagrees (:378-406).
NZ_HD inline void step_begin(const Grid& g, const Lane& s, StepState& q, double dt) {
  q.jr = 1.0;
  q.niter = 0;
  q.pending = 0;
  if (s.sid != q.ds_sid) {  // the structure changed since ds was computed
    q.ds = sqrt(2.0 * g.D[(int)s.sid] * dt);
    q.ds_sid = s.sid;
  }
  const int4 t = g.cell_tab[cell_id(g, s.l[0], s.l[1], s.l[2])];
  q.cs = t.x;
  q.ce = t.y;
  q.rs = t.z;
}

// Draw the pending jump: draw 0 at the start of a step, draw k after the
// k-th sub-step scattered (:403-405, :476-478).
NZ_HD inline void step_draw(StepState& q, uint64_t seed, uint32_t spin, uint32_t step) {
  draw_jump(seed, spin, step, (uint32_t)q.pending, q.ds, q.j);
  q.pending = -1;
}

// x += jc j (sim3d_kernel.cu:460-462), each product rounded before the add.
NZ_HD inline void move(Lane& s, const StepState& q, double jc) {
#pragma clang fp contract(off)
  s.x[0] += jc * q.j[0];
  s.x[1] += jc * q.j[1];
  s.x[2] += jc * q.j[2];
}

// Remaining fraction after a scatter, sqrt(max(0, jr^2 - jc^2)) (:470-471),
// each square rounded before the difference.
NZ_HD inline double scatter_remaining(double jr, double jc) {
#pragma clang fp contract(off)
  return sqrt(MAXd(0.0, jr * jr - jc * jc));
}

// One sub-step of the multi-collision loop (sim3d_kernel.cu:373-487).
// Returns true when the step is complete (:373 loop condition).
NZ_HD inline bool substep(const Grid& g, const ListBufs& lb, Lane& s, StepState& q, uint64_t seed,
                          uint32_t spin, uint32_t step, double dt, bool do_t2) {
  q.niter++;
  double jc = q.jr;  // :376
  double rel[3];
  lane_rel(g, s, rel);
  ScanAcc acc;
  for (int pass = 0; pass < 3; ++pass) {
    int exact_mode = -1;
    ScanIn in;
    scan_in(g, rel, q.j, q.jr, s.sid, in);
    if (!in_safety_f(g, in, (float)q.jr)) {
      if (rebin(g, rel, s.l)) {
        const int4 t = g.cell_tab[cell_id(g, s.l[0], s.l[1], s.l[2])];
        q.cs = t.x;
        q.ce = t.y;
        q.rs = t.z;
        scan_in(g, rel, q.j, q.jr, s.sid, in);
        if (!in_safety_f(g, in, (float)q.jr))
          exact_mode = in_neighbourhood(g, rel, q.j, q.jr, s.l) ? kExactNeighbours : kExactAll;
      } else {
        exact_mode = kExactAll;
      }
    }
    acc_reset(g, acc);
    if (exact_mode < 0 && q.ce - q.cs > kMaxListLen) exact_mode = kExactList;  // dense cell
    if (exact_mode < 0) {
      ScanKeys k;
      keys_reset(g, k);
      scan_list(g, lb, q.cs, q.ce, q.rs, in, k);
      if (k.gcnt > 1) {
        exact_mode = kExactList;
      } else {
        keys_decode(g, q.cs, k, acc);
        if (k.gcnt == 1) {  // the one undecided candidate in FP64
          s.nguard++;
          visit_exact(g, s.x, q.j, s.sid, cand_sphere(g, q.cs + (int)k.gpos - 1), acc);
        }
      }
    }
    if (exact_mode >= 0) {  // rare: the scan in FP64
      s.nguard++;
      scan_exact(g, exact_mode, q.cs, q.ce, s.x, q.j, s.l, s.sid, acc);
    }
    if (acc.sid_new == s.sid) break;
    s.sid = acc.sid_new;  // structure changed: redo with the new id
    if (q.niter == 1) {
      q.ds = sqrt(2.0 * g.D[(int)s.sid] * dt);
      q.ds_sid = s.sid;
      draw_jump(seed, spin, step, opaque_u32(0u), q.ds, q.j);
    }
  }
  // The decisive roots in FP64 (the oracle's arithmetic).  A root whose FP32
  // value exceeds the smaller of jr and the other kind's root by more than
  // FP32 round-off cannot decide jc = min(jr, fin, fout) and is not
  // re-solved; a runner-up within that round-off of the best of its kind is
  // re-solved beside it (the FP32 keys cannot order the two).
  double fin = acc.any_in ? (double)acc.fin : (double)kLarge;
  double fout = (double)acc.fout;
  {
    const float fi = acc.any_in ? acc.fin : kLarge;
    const float lim_in = fminf(acc.fout, (float)q.jr) * (1.0f + kTieRel) + kTieAbs;
    const float lim_out = fminf(fi, (float)q.jr) * (1.0f + kTieRel) + kTieAbs;
    const bool need_in = acc.kin >= 0 && acc.any_in && fi <= lim_in;
    const bool need_out = acc.kout >= 0 && acc.fout <= lim_out;
    const bool need_in2 = need_in && acc.kin2 >= 0 && acc.fin2 >= fi * (1.0f - kTieRel) - kTieAbs;
    const bool need_out2 =
        need_out && acc.kout2 >= 0 && acc.fout2 <= acc.fout * (1.0f + kTieRel) + kTieAbs;
    if (acc.any_in && acc.kin < 0) fin = 0.0;
    // Usually one root per lane: the loop runs as often as the lane of the
    // wavefront with the most near-ties needs.
    int todo = (need_in ? 1 : 0) | (need_in2 ? 2 : 0) | (need_out ? 4 : 0) | (need_out2 ? 8 : 0);
#pragma unroll 1
    while (todo != 0) {
      const int bit = todo & -todo;
      todo &= todo - 1;
      const bool in_root = bit <= 2;
      const int k = bit == 1 ? acc.kin : bit == 2 ? acc.kin2 : bit == 4 ? acc.kout : acc.kout2;
      const double4 c = g.sph[k];
      const double v = root_d(in_root, s.x[0] - c.x, s.x[1] - c.y, s.x[2] - c.z, c.w, q.j[0],
                              q.j[1], q.j[2]);
      if (bit == 1) fin = v;
      else if (bit == 2) fin = MAXd(fin, v);
      else if (bit == 4) fout = v;
      else fout = MINd(fout, v);
    }
  }
  jc = MINd(jc, acc.any_in ? fin : (double)kLarge);  // :457
  jc = MINd(jc, fout);                               // :458
  move(s, q, jc);                                    // :460-462
  const float inv_t2 = g.inv_T2[(int)s.sid];
  if (jc < q.jr) {  // scatter (:464-479)
    q.jr = scatter_remaining(q.jr, jc);
    if (do_t2) s.logsig = fmaf((float)(-jc * jc * dt), inv_t2, s.logsig);
    s.nsub++;
    if (q.jr > (double)kTol && q.niter < kMaxIter)  // the draw is only used by a further sub-step
      q.pending = q.niter;
  } else {  // coherent (:480-486)
    q.jr -= jc;
    if (do_t2) s.logsig = fmaf((float)(-jc * dt), inv_t2, s.logsig);
  }
  return !(q.jr > (double)kTol && q.niter < kMaxIter);
}

// End of a time step: phase after the move and before the wrap (:674-676,
// :880-887), the periodic wrap once per axis with the spins0 shift and the
// phase correction (:489-521, :683-715, :891-942), then the cell holding the
// end point.
// with_m false: no phase moment is kept (its registers are left to the walk).
NZ_HD inline void step_end(const Grid& g, Lane& s, float G, float Ga, bool with_m = true) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (with_m) s.m[a] = fmaf(G, (float)s.x[a], s.m[a]);
    const double La = g.Ld[a];
    int l = s.l[a];
    if (s.x[a] < -La / 2) {
      s.x[a] += La;
      s.w[a] += 1;
      if (with_m) s.m[a] = fmaf(g.L[a], Ga, s.m[a]);
      l += g.ncell[a];
    } else if (s.x[a] > La / 2) {
      s.x[a] -= La;
      s.w[a] -= 1;
      if (with_m) s.m[a] = fmaf(-g.L[a], Ga, s.m[a]);
      l -= g.ncell[a];
    }
    // a step rarely leaves the neighbouring cell: two comparisons, the
    // division only as the fallback
    const double r = s.x[a] - cell_lo(g, a, l);
    if (r < 0.0) --l;
    else if (r >= g.hd[a]) ++l;
    const double r2 = s.x[a] - cell_lo(g, a, l);
    if (l < 0 || l > g.ncell[a] + 1 || r2 < 0.0 || r2 >= g.hd[a]) l = cell_of(g, a, s.x[a]);
    s.l[a] = l;
  }
}

// One whole time step of one spin (the sequential form of the above).
NZ_HD inline void walk_step(const Grid& g, Lane& s, uint64_t seed, uint32_t spin, uint32_t step,
                            double dt, bool do_t2, float G, float Ga) {
  StepState q;
  q.ds_sid = -1.0f;
  step_begin(g, s, q, dt);
  bool done = false;
  while (!done) {
    if (q.pending >= 0) step_draw(q, seed, spin, step);
    done = substep(g, list_bufs(g), s, q, seed, spin, step, dt, do_t2);
  }
  step_end(g, s, G, Ga);
}

// Displacement since the start of the spin (unwrapped), per axis: position
// minus the wrap-shifted start position, as the reference's x - spins0.
NZ_HD inline double lane_disp(const Grid& g, const Lane& s, const double x0[3], int a) {
#pragma clang fp contract(off)
  return s.x[a] - (x0[a] + (double)s.w[a] * g.Ld[a]);
}

// Absolute position -> lane state with the structure id of the point.
NZ_HD inline void lane_init(const Grid& g, const float pos[3], Lane& s) {
#pragma clang fp contract(off)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.x[a] = (double)pos[a];
    int c = (int)floor((s.x[a] + 0.5 * g.Ld[a]) / g.hd[a]);
    c = c < 0 ? 0 : (c > g.ncell[a] - 1 ? g.ncell[a] - 1 : c);
    s.l[a] = c + 1;
    s.w[a] = 0;
    s.m[a] = 0.0f;
  }
  // the cell list holds every sphere that can contain a point of the cell;
  // decided in FP64 with the reference's inside test
  const int cid = cell_id(g, s.l[0], s.l[1], s.l[2]);
  float sid = (float)g.nstruct;
  for (int k = g.cell_start[cid]; k < g.cell_start[cid + 1]; ++k) {
    const int idx = cand_sphere(g, k);
    const double4 c = g.sph[idx];
    const double dx = s.x[0] - c.x, dy = s.x[1] - c.y, dz = s.x[2] - c.z;
    const float sc = (float)g.sph_sid[idx];
    if ((dx * dx + dy * dy + dz * dz - c.w * c.w) < 0 && sc < sid) sid = sc;
  }
  s.sid = sid;
  s.logsig = 0.0f;
  s.nsub = 0;
  s.nguard = 0;
}

// Wrapped absolute position and the start position shifted by every wrap
// (the reference's spins / spins0 after its wrap rule).
NZ_HD inline void lane_outputs(const Grid& g, const Lane& s, const double x0[3], float pos[3],
                               float pos0[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    pos[a] = (float)s.x[a];
    pos0[a] = (float)(x0[a] + (double)s.w[a] * g.Ld[a]);
  }
}

}  // namespace nz
