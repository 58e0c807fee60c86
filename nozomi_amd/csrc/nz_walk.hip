// The walker: one spin per lane, many steps per launch.
//
// Algorithm of the reference walker (sim3d_kernel.cu:340-522, with the phase
// additions of :524-717 and :719-943), restructured for gfx950:
//  * positions live in registers as (cell index, cell-relative float) across
//    all steps of a launch; single precision then resolves ~1e-7 um instead
//    of the ~2e-6 um of absolute coordinates in a 60 um box;
//  * the structure pass and the collision pass over the candidate list are
//    fused into one pass that assumes the structure id of the previous
//    sub-step and re-runs only when the id changed (ids only change through
//    numerical leakage), which is exact with respect to the two-pass order;
//  * every sub-step checks that its segment stays inside the safety box of
//    its cell list (cell grown by the list margin) and re-bins or falls back
//    to all spheres when it does not, so no jump can miss a sphere;
//  * the T2 weight is kept as a log (one FMA per sub-step instead of an exp);
//  * phase is kept as the 3-vector moment sum_t g_t x_t: every direction of
//    the reference's multi-direction kernel is a projection of it, so the
//    per-step cost is independent of the number of directions.
#include <hip/hip_runtime.h>

#include "nz_internal.h"
#include "nz_rng.h"

namespace nz {

// the reference's comparison macros (sim3d_kernel.cu:7-8), NaN behaviour kept
__device__ __forceinline__ float MINr(float a, float b) { return (a > b) ? b : a; }
__device__ __forceinline__ float MAXr(float a, float b) { return (a < b) ? b : a; }

struct WalkArgs {
  Grid g;
  nz_spin_state st;
  uint64_t seed;
  int64_t step0;
  int32_t nsteps;
  float dt;
  const float* gph;
  const float* garea;
  double* stepmom;
  uint32_t flags;
};

struct ScanAcc {
  float sid_new;  // smallest id of a containing sphere (box id if none)
  float fin;      // furthest exit fraction among containing spheres
  float fout;     // closest entry fraction among lower-id spheres
  bool any_in;    // inside some relevant sphere
};

// One candidate sphere: the structure test (sim3d_kernel.cu:380-398) and the
// collision test (:414-455) fused.  A = (cx, cy, cz, sid) relative to the
// lane's cell corner, B = (r^2, (r-tol)^2, (r+tol)^2, r).  The two root
// formulas are those of findStepSizeToSphereInsideTol / ...OutsideTol
// (:77-160) evaluated on one shared quadratic.
__device__ __forceinline__ void visit(float px, float py, float pz, float jx, float jy, float jz,
                                      float a, float inv_a, float sid, const float4 A,
                                      const float4 B, ScanAcc& acc) {
  const float dx = px - A.x, dy = py - A.y, dz = pz - A.z;
  const float dd = dx * dx + dy * dy + dz * dz;
  const bool inside = (dd - B.x) < 0.0f;                       // isInsideSphere < 0
  if (inside && A.w < acc.sid_new) acc.sid_new = A.w;           // :384-396
  const bool relevant = A.w <= sid;                             // :418-419
  const float b = 2.0f * (jx * dx + jy * dy + jz * dz);
  const float c = dd - (inside ? B.y : B.z);
  const float disc = b * b - 4.0f * a * c;
  const float sq = __builtin_sqrtf(fmaxf(disc, 0.0f));
  const float q = (b < 0.0f) ? -0.5f * (b - sq) : -0.5f * (b + sq);
  const float fpos = q * inv_a;
  const float fneg = c * __builtin_amdgcn_rcpf(q);
  if (relevant && inside) {                                     // :428-440
    const float f = (disc < 0.0f) ? 0.0f : MAXr(MAXr(fpos, fneg), 0.0f);
    acc.any_in = true;
    if (acc.fin < f) acc.fin = f;
  } else if (relevant && A.w < sid) {                           // :441-454
    float f;
    if (disc < 0.0f) f = kLarge;
    else if (fpos > 0.0f && fneg > 0.0f) f = MINr(fpos, fneg);
    else if ((fpos < 0.0f && fneg < 0.0f) || (b > 0.0f)) f = kLarge;
    else f = 0.0f;
    if (f < acc.fout) acc.fout = f;
  }
}

__device__ __forceinline__ int cell_id(const Grid& g, int lx, int ly, int lz) {
  return (lz * g.nloc[1] + ly) * g.nloc[0] + lx;
}

// Segment p -> p + jr*j inside the safety box [-m, h+m]^3 of the lane's cell?
__device__ __forceinline__ bool in_safety(const Grid& g, const float p[3], const float j[3],
                                          float jr) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float e = p[a] + jr * j[a];
    const float lo = -g.margin, hi = g.h[a] + g.margin;
    ok = ok && p[a] >= lo && p[a] <= hi && e >= lo && e <= hi;
  }
  return ok;
}

// Move the lane's frame to the cell holding p (in-step re-bin).  Returns
// false when that cell is outside the grid's halo.
__device__ __forceinline__ bool rebin(const Grid& g, float p[3], int l[3], int u[3]) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int k = (int)floorf(p[a] * g.inv_h[a]);
    if (k != 0) {
      const int nl = l[a] + k;
      if (nl < 0 || nl > g.ncell[a] + 1) {
        ok = false;
      } else {
        l[a] = nl;
        u[a] += k;
        p[a] = fmaf(-(float)k, g.h[a], p[a]);
      }
    }
  }
  return ok;
}

__device__ __forceinline__ void scan_list(const Grid& g, int cs, int ce, const float p[3],
                                          const float j[3], float a, float inv_a, float sid,
                                          ScanAcc& acc) {
  for (int k = cs; k < ce; ++k) {
    const float4 A = g.cand_a[k];
    const float4 B = g.cand_b[k];
    visit(p[0], p[1], p[2], j[0], j[1], j[2], a, inv_a, sid, A, B, acc);
  }
}

// Every sphere, converted to the lane's frame in double precision.
__device__ __noinline__ void scan_all(const Grid& g, const int l[3], const float p[3],
                                      const float j[3], float a, float inv_a, float sid,
                                      ScanAcc& acc) {
  double lo[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    lo[ax] = -0.5 * (double)g.L[ax] + (double)(l[ax] - 1) * ((double)g.L[ax] / g.ncell[ax]);
  const double tol = (double)kTol;
  for (int64_t k = 0; k < g.nspheres; ++k) {
    const double4 s = g.sph[k];
    const float4 A = make_float4((float)(s.x - lo[0]), (float)(s.y - lo[1]), (float)(s.z - lo[2]),
                                 (float)g.sph_sid[k]);
    const float4 B = make_float4((float)(s.w * s.w), (float)((s.w - tol) * (s.w - tol)),
                                 (float)((s.w + tol) * (s.w + tol)), (float)s.w);
    visit(p[0], p[1], p[2], j[0], j[1], j[2], a, inv_a, sid, A, B, acc);
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(256) walk_kernel(const WalkArgs A) {
  const Grid& g = A.g;
  const nz_spin_state& st = A.st;
  const int64_t n = st.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < n;
  const int64_t ii = active ? i : 0;

  int u[3], l[3];
  float p[3];
  float logsig = 0.0f, m[3] = {0.0f, 0.0f, 0.0f};
  float sid = (float)g.nstruct;
  int nsub = 0;
  int u0[3];
  float p0[3];
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      u[a] = st.ucell[a * n + ii];
      p[a] = st.rel[a * n + ii];
      u0[a] = st.ucell0[a * n + ii];
      p0[a] = st.rel0[a * n + ii];
      int w = u[a] % g.ncell[a];
      if (w < 0) w += g.ncell[a];
      l[a] = w + 1;
    }
    logsig = st.logsig[ii];
    sid = (float)st.sid[ii];
    if (A.gph && st.mom) {
      m[0] = st.mom[ii];
      m[1] = st.mom[n + ii];
      m[2] = st.mom[2 * n + ii];
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      u[a] = 0; l[a] = 1; p[a] = 0.0f; u0[a] = 0; p0[a] = 0.0f;
    }
  }
  const bool do_t2 = (A.flags & NZ_WALK_T2) != 0;
  const uint32_t spin = (uint32_t)ii;

  for (int t = 0; t < A.nsteps; ++t) {
    const uint32_t step = (uint32_t)(A.step0 + t);
    float jr = active ? 1.0f : 0.0f;  // :364
    int niter = 0;
    float ds = 0.0f, j[3] = {0.0f, 0.0f, 0.0f};
    int cid = cell_id(g, l[0], l[1], l[2]);
    int cs = g.cell_start[cid], ce = g.cell_start[cid + 1];
    while (jr > kTol && niter < kMaxIter) {  // :373
      niter++;
      float jc = jr;
      if (niter == 1) {  // :400-406
        ds = sqrtf(2.0f * g.D[(int)sid] * A.dt);
        float n0, n1, n2;
        normals3(A.seed, kStreamWalk, spin, step, 0u, n0, n1, n2);
        j[0] = ds * n0; j[1] = ds * n1; j[2] = ds * n2;
      }
      ScanAcc acc;
      for (int pass = 0; pass < 3; ++pass) {
        bool brute = false;
        if (!in_safety(g, p, j, jr)) {
          if (rebin(g, p, l, u)) {
            cid = cell_id(g, l[0], l[1], l[2]);
            cs = g.cell_start[cid];
            ce = g.cell_start[cid + 1];
            brute = !in_safety(g, p, j, jr);
          } else {
            brute = true;
          }
        }
        const float a = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
        const float inv_a = 1.0f / a;
        acc.sid_new = (float)g.nstruct;
        acc.fin = 0.0f;
        acc.fout = kLarge;
        acc.any_in = false;
        if (!brute) scan_list(g, cs, ce, p, j, a, inv_a, sid, acc);
        else scan_all(g, l, p, j, a, inv_a, sid, acc);
        if (acc.sid_new == sid) break;
        sid = acc.sid_new;  // structure changed: redo with the new id
        if (niter == 1) {
          ds = sqrtf(2.0f * g.D[(int)sid] * A.dt);
          float n0, n1, n2;
          normals3(A.seed, kStreamWalk, spin, step, 0u, n0, n1, n2);
          j[0] = ds * n0; j[1] = ds * n1; j[2] = ds * n2;
        }
      }
      jc = MINr(jc, acc.any_in ? acc.fin : kLarge);  // :457
      jc = MINr(jc, acc.fout);                       // :458
      p[0] = fmaf(jc, j[0], p[0]);                   // :460-462
      p[1] = fmaf(jc, j[1], p[1]);
      p[2] = fmaf(jc, j[2], p[2]);
      const float inv_t2 = 1.0f / g.T2[(int)sid];
      if (jc < jr) {  // scatter (:464-479)
        jr = sqrtf(MAXr(0.0f, jr * jr - jc * jc));
        if (do_t2) logsig = fmaf(-jc * jc * A.dt, inv_t2, logsig);
        nsub++;
        if (jr > kTol && niter < kMaxIter) {  // the draw is only used by a further sub-step
          float n0, n1, n2;
          normals3(A.seed, kStreamWalk, spin, step, (uint32_t)niter, n0, n1, n2);
          j[0] = ds * n0; j[1] = ds * n1; j[2] = ds * n2;
        }
      } else {  // coherent (:480-486)
        jr -= jc;
        if (do_t2) logsig = fmaf(-jc * A.dt, inv_t2, logsig);
      }
    }
    // phase after the move, before the wrap (:674-676, :880-887)
    float G = 0.0f, Ga = 0.0f;
    if (A.gph) {
      G = A.gph[t];
      Ga = A.garea[t];
    }
    // re-base to the cell holding the end point, then wrap (:489-521)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float x = fmaf((float)(l[a] - 1), g.h[a], -0.5f * g.L[a]) + p[a];
      m[a] = fmaf(G, x, m[a]);
      int k = (int)floorf(p[a] * g.inv_h[a]);
      float r = (k != 0) ? fmaf(-(float)k, g.h[a], p[a]) : p[a];
      if (r < 0.0f) { r += g.h[a]; --k; }
      if (r >= g.h[a]) { r -= g.h[a]; ++k; }
      int nl = l[a] + k;
      if (nl == g.ncell[a] + 1 && r == 0.0f) { --nl; --k; r = g.h[a]; }  // x == L/2 stays
      p[a] = r;
      u[a] += k;
      // wrap: x > L/2  <=>  nl > ncell;  x < -L/2  <=>  nl < 1.  w = net
      // number of +L wraps; each one shifts the phase by -L*G_area (:891-942)
      int w = 0;
      if (nl > g.ncell[a]) w = (nl - 1) / g.ncell[a];
      else if (nl < 1) w = -((g.ncell[a] - nl) / g.ncell[a]);
      if (w != 0) {
        nl -= w * g.ncell[a];
        m[a] = fmaf(-(float)w * g.L[a], Ga, m[a]);
      }
      l[a] = nl;
    }
    if (A.stepmom) {  // moment sums (sim3d_kernel.cu:191-237), summed in FP64
      float v[6];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float d = active ? fmaf((float)(u[a] - u0[a]), g.h[a], p[a] - p0[a]) : 0.0f;
        v[a] = d * d;
        v[3 + a] = v[a] * v[a];
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const float s = wave_sum(v[q]);
        if ((threadIdx.x & 63) == 0) atomicAdd(&A.stepmom[(int64_t)t * 6 + q], (double)s);
      }
    }
  }

  if (!active) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.ucell[a * n + i] = u[a];
    st.rel[a * n + i] = p[a];
    const double hd = (double)g.L[a] / g.ncell[a];
    // wrapped absolute position and the start position shifted by every wrap
    const int wraps = (u[a] >= 0) ? u[a] / g.ncell[a] : -((-u[a] + g.ncell[a] - 1) / g.ncell[a]);
    st.pos[a * n + i] = (float)(-0.5 * (double)g.L[a] + (double)(l[a] - 1) * hd + (double)p[a]);
    st.pos0[a * n + i] = (float)(-0.5 * (double)g.L[a] + (double)u0[a] * hd + (double)p0[a] -
                                 (double)wraps * (double)g.L[a]);
  }
  st.logsig[i] = logsig;
  st.sig[i] = expf(logsig);
  st.sid[i] = (int32_t)sid;
  if (A.gph && st.mom) {
    st.mom[i] = m[0];
    st.mom[n + i] = m[1];
    st.mom[2 * n + i] = m[2];
  }
  if (st.nsub) st.nsub[i] += nsub;
}

// Absolute positions -> resident state.
__global__ void __launch_bounds__(256) state_init_kernel(const Grid g, const nz_spin_state st) {
  const int64_t n = st.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int l[3];
  float p[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double hd = (double)g.L[a] / g.ncell[a];
    const double x = (double)st.pos[a * n + i];
    int c = (int)floor((x + 0.5 * (double)g.L[a]) / hd);
    c = c < 0 ? 0 : (c > g.ncell[a] - 1 ? g.ncell[a] - 1 : c);
    const float r = (float)(x - (-0.5 * (double)g.L[a] + (double)c * hd));
    l[a] = c + 1;
    p[a] = r;
    st.ucell[a * n + i] = c;
    st.ucell0[a * n + i] = c;
    st.rel[a * n + i] = r;
    st.rel0[a * n + i] = r;
    st.pos0[a * n + i] = st.pos[a * n + i];
  }
  // structure id of the start point: the cell list holds every sphere that
  // can contain a point of the cell
  const int cid = cell_id(g, l[0], l[1], l[2]);
  float sid = (float)g.nstruct;
  for (int k = g.cell_start[cid]; k < g.cell_start[cid + 1]; ++k) {
    const float4 A = g.cand_a[k];
    const float4 B = g.cand_b[k];
    const float dx = p[0] - A.x, dy = p[1] - A.y, dz = p[2] - A.z;
    if ((dx * dx + dy * dy + dz * dz - B.x) < 0.0f && A.w < sid) sid = A.w;
  }
  st.sid[i] = (int32_t)sid;
  st.logsig[i] = 0.0f;
  st.sig[i] = 1.0f;
  if (st.mom) {
    st.mom[i] = 0.0f;
    st.mom[n + i] = 0.0f;
    st.mom[2 * n + i] = 0.0f;
  }
  if (st.nsub) st.nsub[i] = 0;
}

__global__ void __launch_bounds__(256) disp_moments_kernel(const Grid g, const nz_spin_state st,
                                                           double* out) {
  const int64_t n = st.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float d = fmaf((float)(st.ucell[a * n + i] - st.ucell0[a * n + i]), g.h[a],
                           st.rel[a * n + i] - st.rel0[a * n + i]);
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

__global__ void zero6(double* p) {
  if (threadIdx.x < 6) p[threadIdx.x] = 0.0;
}

static int check_state(const nz_spin_state* st, const char* who) {
  if (!st) return fail(-1, std::string(who) + ": state is NULL");
  if (st->n <= 0 || st->n > (int64_t)INT32_MAX)
    return fail(-1, std::string(who) + ": spin count must be in [1, 2^31)");
  if (!st->pos || !st->pos0 || !st->sig || !st->ucell || !st->rel || !st->ucell0 || !st->rel0 ||
      !st->logsig || !st->sid)
    return fail(-1, std::string(who) + ": a required state buffer is NULL");
  return 0;
}

}  // namespace nz

using namespace nz;

extern "C" {

int nz_state_init(nz_substrate_t sub, const nz_spin_state* st, void* stream) {
  if (!sub) return fail(-1, "nz_state_init: substrate is NULL");
  if (int rc = check_state(st, "nz_state_init")) return rc;
  const Substrate* s = (const Substrate*)sub;
  const int64_t blocks = (st->n + 255) / 256;
  hipLaunchKernelGGL(state_init_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     s->g, *st);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_displacement_moments(nz_substrate_t sub, const nz_spin_state* st, double* out, void* stream) {
  if (!sub) return fail(-1, "nz_displacement_moments: substrate is NULL");
  if (int rc = check_state(st, "nz_displacement_moments")) return rc;
  if (!out) return fail(-1, "nz_displacement_moments: out is NULL");
  const Substrate* s = (const Substrate*)sub;
  hipLaunchKernelGGL(zero6, dim3(1), dim3(64), 0, (hipStream_t)stream, out);
  const int64_t blocks = (st->n + 255) / 256;
  hipLaunchKernelGGL(disp_moments_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     s->g, *st, out);
  NZ_HIP(hipGetLastError());
  return 0;
}

int nz_walk(nz_substrate_t sub, const nz_spin_state* st, uint64_t seed, int64_t step0,
            int32_t nsteps, float dt, const float* g, const float* garea, double* stepmom,
            uint32_t flags, void* stream) {
  if (!sub) return fail(-1, "nz_walk: substrate is NULL");
  if (int rc = check_state(st, "nz_walk")) return rc;
  if (nsteps < 0) return fail(-1, "nz_walk: nsteps < 0");
  if (!(dt >= 0.0f)) return fail(-1, "nz_walk: dt must be >= 0");
  if ((g == nullptr) != (garea == nullptr)) return fail(-1, "nz_walk: pass both g and garea or neither");
  if (g && !st->mom) return fail(-1, "nz_walk: phase coefficients given but state.mom is NULL");
  if (step0 < 0 || step0 + nsteps > (int64_t)UINT32_MAX) return fail(-1, "nz_walk: step index out of range");
  if (nsteps == 0) return 0;
  const Substrate* s = (const Substrate*)sub;
  WalkArgs a;
  a.g = s->g;
  a.st = *st;
  a.seed = seed;
  a.step0 = step0;
  a.nsteps = nsteps;
  a.dt = dt;
  a.gph = g;
  a.garea = garea;
  a.stepmom = stepmom;
  a.flags = flags;
  const int64_t blocks = (st->n + 255) / 256;
  hipLaunchKernelGGL(walk_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, a);
  NZ_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
