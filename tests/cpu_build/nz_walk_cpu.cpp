// CPU build of the walker's per-spin core (nozomi_amd/csrc/nz_walk_core.h and
// nz_grid_build.h), for tests that check the kernel logic against the FP64
// oracle without a GPU.  Test infrastructure only.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../nozomi_amd/csrc/nz_grid_build.h"
#include "../../nozomi_amd/csrc/nz_walk_core.h"

namespace nz {
// error helpers referenced by nz_internal.h (the library defines them in
// nz_substrate.hip)
static std::string g_cpu_err;
void set_error(const std::string& msg) { g_cpu_err = msg; }
int fail(int code, const std::string& msg) { set_error(msg); return code; }
int hip_fail(hipError_t e, const char* what) { set_error(what); return (int)e; }
}  // namespace nz

using namespace nz;

struct CpuWalker {
  HostGrid hg;
  std::vector<Lane> lanes;
  std::vector<double> x0;
};

extern "C" {

void* nzcpu_create(const double* spheres, int64_t S, int32_t nstruct, const double* D,
                   const double* T2, double Lx, double Ly, double Lz, double cell, double margin) {
  CpuWalker* w = new CpuWalker();
  const double L[3] = {Lx, Ly, Lz};
  std::string err;
  if (build_grid_host(spheres, S, nstruct, D, T2, L, cell, margin, w->hg, err) != 0) {
    delete w;
    return nullptr;
  }
  w->hg.bind_host();
  return w;
}

void nzcpu_destroy(void* h) { delete (CpuWalker*)h; }

// The two-phase candidate scan against the visit of every candidate, for n
// probes: positions pos [3, n] (absolute), jump vectors jump [3, n], remaining
// fractions jr [n], assumed structure ids sid [n].  Each probe scans the list
// of the cell holding its position.  out[4 i + k] = 1 where
//   k = 0: the exit keys (kin) differ;
//   k = 1: the entry keys (kout) differ although one of them is a root that
//          can decide the sub-step (not beyond jr 1.0001 + 1e-6, the bound of
//          the FP64 re-solve); entries beyond it cannot, and the two-phase scan
//          skips the obstacles they belong to;
//   k = 2: the structure ids differ;
//   k = 3: the guards disagree beyond the allowed difference (the two-phase
//          scan skips out-of-reach guards: it may count fewer, never more, and
//          a single guard must be the same one).
void nzcpu_scan_compare(void* h, int n, const double* pos, const double* jump, const double* jr,
                        const float* sid, int32_t* out) {
  const Grid& g = ((CpuWalker*)h)->hg.g;
  for (int i = 0; i < n; ++i) {
    Lane s;
    const float p[3] = {(float)pos[i], (float)pos[n + i], (float)pos[2 * n + i]};
    lane_init(g, p, s);
    StepState q;
    q.jr = jr[i];
    for (int a = 0; a < 3; ++a) q.j[a] = jump[a * n + i];
    const int cid = cell_id(g, s.l[0], s.l[1], s.l[2]);
    const int cs = g.cell_start[cid], ce = g.cell_start[cid + 1];
    double rel[3];
    lane_rel(g, s, rel);
    ScanIn in;
    scan_in(g, rel, q.j, q.jr, sid[i], in);
    ScanKeys k2, k1;
    keys_reset(g, k2);
    keys_reset(g, k1);
    scan_list(g, list_bufs(g), cs, ce, g.reach_start[cid], in, k2);
    scan_all(g.cand + 2 * (int64_t)cs, ce - cs, in, k1);
    out[4 * i + 0] = k2.kin != k1.kin;
    const float lim = (float)q.jr * 1.0001f + 1e-6f;
    const float f1 = (k1.kout != 0xffffffffu) ? u2f(k1.kout & kKeyMask) : kLarge;
    const float f2 = (k2.kout != 0xffffffffu) ? u2f(k2.kout & kKeyMask) : kLarge;
    out[4 * i + 1] = (k2.kout != k1.kout) && (f1 <= lim || f2 <= lim);
    out[4 * i + 2] = k2.sid_new != k1.sid_new;
    out[4 * i + 3] = (k2.gcnt > k1.gcnt) || (k1.gcnt == 1 && k2.gcnt == 1 && k1.gpos != k2.gpos);
  }
}

// The reach bound against the FP64 distance from the segment p -> p + jr j
// to each candidate's centre.  out[3 i] = candidates whose R-ball (R = r +
// tol) the segment touches but the FP32 reach test rejects (must be 0),
// out[3 i + 1] = candidates the segment touches, out[3 i + 2] = candidates
// the reach test passes.
void nzcpu_reach_check(void* h, int n, const double* pos, const double* jump, const double* jr,
                       int32_t* out) {
  const Grid& g = ((CpuWalker*)h)->hg.g;
  for (int i = 0; i < n; ++i) {
    Lane s;
    const float p[3] = {(float)pos[i], (float)pos[n + i], (float)pos[2 * n + i]};
    lane_init(g, p, s);
    double j[3];
    for (int a = 0; a < 3; ++a) j[a] = jump[a * n + i];
    const int cid = cell_id(g, s.l[0], s.l[1], s.l[2]);
    double rel[3];
    lane_rel(g, s, rel);
    ScanIn in;
    scan_in(g, rel, j, jr[i], 0.0f, in);
    int32_t miss = 0, touch = 0, pass = 0;
    for (int e = g.cell_start[cid]; e < g.cell_start[cid + 1]; ++e) {
      const int k = cand_sphere(g, e);
      const double4 c = g.sph[k];
      // closest point of the segment to the centre
      double d[3], jj = 0.0, dj = 0.0;
      for (int a = 0; a < 3; ++a) {
        d[a] = c.x * (a == 0) + c.y * (a == 1) + c.z * (a == 2) - s.x[a];
        jj += j[a] * j[a] * jr[i] * jr[i];
        dj += d[a] * j[a] * jr[i];
      }
      const double t = jj > 0 ? std::min(1.0, std::max(0.0, dj / jj)) : 0.0;
      double dist2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double v = d[a] - t * j[a] * jr[i];
        dist2 += v * v;
      }
      const double R = c.w + (double)kTol;
      const bool touches = dist2 <= R * R;
      const uint2 w = g.reach[g.reach_start[cid] + (e - g.cell_start[cid])];
      const bool passes = reach_entry(in, w.x, w.y);
      touch += touches;
      pass += passes;
      miss += touches && !passes;
    }
    out[3 * i] = miss;
    out[3 * i + 1] = touch;
    out[3 * i + 2] = pass;
  }
}

// The reach test's parameters: lambda, sigma, sigma (1 + 1/lambda), eps.
void nzcpu_reach_params(void* h, double* out) {
  const Grid& g = ((CpuWalker*)h)->hg.g;
  out[0] = g.reach_lambda;
  out[1] = g.reach_sigma;
  out[2] = g.reach_lam1;
  out[3] = g.reach_eps;
}

// The host builder's reach entries: sizes[0] = entries (padding included),
// sizes[1] = cells; reach_start [cells], reach [2 entries] words.
void nzcpu_reach(void* h, int32_t* reach_start, uint32_t* reach, int64_t* sizes) {
  const HostGrid& hg = ((CpuWalker*)h)->hg;
  sizes[0] = (int64_t)hg.reach.size();
  sizes[1] = (int64_t)hg.reach_start.size();
  if (reach_start) std::memcpy(reach_start, hg.reach_start.data(), sizeof(int32_t) * sizes[1]);
  if (reach) std::memcpy(reach, hg.reach.data(), sizeof(uint2) * sizes[0]);
}

// the walk's normals (host build of nz::normals3), n draws from draw0
void nzcpu_normals(uint64_t seed, uint32_t stream, uint32_t spin, uint32_t step, uint32_t draw0,
                   int n, float* out) {
  for (int k = 0; k < n; ++k)
    nz::normals3(seed, stream, spin, step, draw0 + (uint32_t)k, out[3 * k], out[3 * k + 1],
                 out[3 * k + 2]);
}

void nzcpu_info(void* h, double* out /* ncell[3], cell[3], margin, ncand_mean_box */) {
  const nz_substrate_info_t& inf = ((CpuWalker*)h)->hg.info;
  for (int a = 0; a < 3; ++a) { out[a] = inf.ncell[a]; out[3 + a] = inf.cell[a]; }
  out[6] = inf.margin;
  out[7] = inf.ncand_mean_box;
}

// pos: float [3, N] absolute start positions
void nzcpu_init(void* h, int N, const float* pos) {
  CpuWalker* w = (CpuWalker*)h;
  w->lanes.resize(N);
  w->x0.resize(3 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    const float p[3] = {pos[i], pos[N + i], pos[2 * N + i]};
    lane_init(w->hg.g, p, w->lanes[i]);
    for (int a = 0; a < 3; ++a) w->x0[3 * i + a] = w->lanes[i].x[a];
  }
}

// Walk every spin nsteps; outputs absolute pos/pos0 [3,N], nsub [N] (total),
// mom [3,N] (when g given), logsig [N].
void nzcpu_walk(void* h, uint64_t seed, int64_t step0, int nsteps, double dt, const float* g,
                const float* garea, int do_t2, float* pos, float* pos0, int32_t* nsub, float* mom,
                float* logsig, int32_t* nguard) {
  CpuWalker* w = (CpuWalker*)h;
  const int N = (int)w->lanes.size();
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < N; ++i) {
    Lane& s = w->lanes[i];
    for (int t = 0; t < nsteps; ++t) {
      const float G = g ? g[t] : 0.0f, Ga = g ? garea[t] : 0.0f;
      walk_step(w->hg.g, s, seed, (uint32_t)i, (uint32_t)(step0 + t), dt, do_t2 != 0, G, Ga);
    }
    float pp[3], pp0[3];
    lane_outputs(w->hg.g, s, &w->x0[3 * i], pp, pp0);
    for (int a = 0; a < 3; ++a) {
      pos[a * N + i] = pp[a];
      pos0[a * N + i] = pp0[a];
      if (mom) mom[a * N + i] = s.m[a];
    }
    if (nsub) nsub[i] = s.nsub;
    if (logsig) logsig[i] = s.logsig;
    if (nguard) nguard[i] = s.nguard;
  }
}

// The host builder's grid (same layout as nz_substrate_export_grid); sizes
// from nzcpu_info: start [nloc^3 + 1], rank [ncell^3], cand [8 (ntotal + 1)].
void nzcpu_grid(void* h, int32_t* start, int32_t* rank, float* cand, int64_t* sizes) {
  const HostGrid& hg = ((CpuWalker*)h)->hg;
  sizes[0] = (int64_t)hg.start.size();
  sizes[1] = (int64_t)hg.rank.size();
  sizes[2] = (int64_t)hg.cand.size() * 4;
  if (start) std::memcpy(start, hg.start.data(), sizeof(int32_t) * hg.start.size());
  if (rank) std::memcpy(rank, hg.rank.data(), sizeof(int32_t) * hg.rank.size());
  if (cand) std::memcpy(cand, hg.cand.data(), sizeof(float4) * hg.cand.size());
}

// FP64 absolute positions (wrapped) and unwrapped displacements [3, N].
void nzcpu_state_f64(void* h, double* pos, double* disp) {
  CpuWalker* w = (CpuWalker*)h;
  const Grid& g = w->hg.g;
  const int N = (int)w->lanes.size();
  for (int i = 0; i < N; ++i) {
    const Lane& s = w->lanes[i];
    for (int a = 0; a < 3; ++a) {
      pos[a * N + i] = s.x[a];
      disp[a * N + i] = lane_disp(g, s, &w->x0[3 * i], a);
    }
  }
}

}  // extern "C"
