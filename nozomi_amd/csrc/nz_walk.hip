// The walker kernels: one spin per lane, many steps per launch (the per-spin
// logic is in nz_walk_core.h).
#include <hip/hip_runtime.h>

#include "nz_walk_core.h"

namespace nz {

struct WalkArgs {
  Grid g;
  nz_spin_state st;
  uint64_t seed;
  int64_t step0;
  int32_t nsteps;
  double dt;
  const float* gph;
  const float* garea;
  double* stepmom;
  uint32_t flags;
};

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

  Lane s;
  int u0[3];
  double p0[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s.u[a] = active ? st.ucell[a * n + ii] : 0;
    s.p[a] = active ? st.rel[a * n + ii] : 0.0;
    u0[a] = active ? st.ucell0[a * n + ii] : 0;
    p0[a] = active ? st.rel0[a * n + ii] : 0.0;
    s.m[a] = (active && A.gph && st.mom) ? st.mom[a * n + ii] : 0.0f;
  }
  lane_load_cells(g, s);
  s.logsig = active ? st.logsig[ii] : 0.0f;
  s.sid = active ? (float)st.sid[ii] : (float)g.nstruct;
  s.nsub = 0;
  s.nguard = 0;
  const bool do_t2 = (A.flags & NZ_WALK_T2) != 0;
  const uint32_t spin = (uint32_t)ii;

  for (int t = 0; t < A.nsteps; ++t) {
    const float G = A.gph ? A.gph[t] : 0.0f;
    const float Ga = A.gph ? A.garea[t] : 0.0f;
    walk_step(g, s, A.seed, spin, (uint32_t)(A.step0 + t), A.dt, do_t2, G, Ga, active);
    if (A.stepmom) {  // moment sums (sim3d_kernel.cu:191-237), summed in FP64
      float v[6];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float d = active ? (float)lane_disp(g, s, u0, p0, a) : 0.0f;
        v[a] = d * d;
        v[3 + a] = v[a] * v[a];
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const float sum = wave_sum(v[q]);
        if ((threadIdx.x & 63) == 0) atomicAdd(&A.stepmom[(int64_t)t * 6 + q], (double)sum);
      }
    }
  }

  if (!active) return;
  float pos[3], pos0[3];
  lane_outputs(g, s, u0, p0, pos, pos0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.ucell[a * n + i] = s.u[a];
    st.rel[a * n + i] = s.p[a];
    st.pos[a * n + i] = pos[a];
    st.pos0[a * n + i] = pos0[a];
  }
  st.logsig[i] = s.logsig;
  st.sig[i] = expf(s.logsig);
  st.sid[i] = (int32_t)s.sid;
  if (A.gph && st.mom) {
    st.mom[i] = s.m[0];
    st.mom[n + i] = s.m[1];
    st.mom[2 * n + i] = s.m[2];
  }
  if (st.nsub) st.nsub[i] += s.nsub;
  if (st.nguard) st.nguard[i] += s.nguard;
}

// Absolute positions -> resident state.
__global__ void __launch_bounds__(256) state_init_kernel(const Grid g, const nz_spin_state st) {
  const int64_t n = st.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Lane s;
  const float pos[3] = {st.pos[i], st.pos[n + i], st.pos[2 * n + i]};
  lane_init(g, pos, s);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    st.ucell[a * n + i] = s.u[a];
    st.ucell0[a * n + i] = s.u[a];
    st.rel[a * n + i] = s.p[a];
    st.rel0[a * n + i] = s.p[a];
    st.pos0[a * n + i] = pos[a];
    if (st.mom) st.mom[a * n + i] = 0.0f;
  }
  st.sid[i] = (int32_t)s.sid;
  st.logsig[i] = 0.0f;
  st.sig[i] = 1.0f;
  if (st.nsub) st.nsub[i] = 0;
  if (st.nguard) st.nguard[i] = 0;
}

__global__ void __launch_bounds__(256) disp_moments_kernel(const Grid g, const nz_spin_state st,
                                                           double* out) {
  const int64_t n = st.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float d = (float)((double)(st.ucell[a * n + i] - st.ucell0[a * n + i]) * g.hd[a] +
                              (st.rel[a * n + i] - st.rel0[a * n + i]));  // lane_disp
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
            int32_t nsteps, double dt, const float* g, const float* garea, double* stepmom,
            uint32_t flags, void* stream) {
  if (!sub) return fail(-1, "nz_walk: substrate is NULL");
  if (int rc = check_state(st, "nz_walk")) return rc;
  if (nsteps < 0) return fail(-1, "nz_walk: nsteps < 0");
  if (!(dt >= 0.0)) return fail(-1, "nz_walk: dt must be >= 0");
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
