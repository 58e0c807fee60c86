// FP64 CPU oracle for the diffusion-MRI random walker.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness checker for the HIP
// walker in nozomi_amd/csrc: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.  The product path never links it.
//
// It restates, in double precision, the algorithm of the reference's CUDA
// walker (simulation_toolkit/simulation_engine/sim3d_kernel.cu) and of the
// host code that drives it (diffsim3d.py, geometry.py).  Every function names
// the reference lines it follows.  Random numbers come from the project's own
// counter-based stream (Philox4x32-10 keyed on (seed, spin, step, draw)); the
// reference's cuRAND XORWOW stream is not reproduced.
//
// Pinning (what checks each function against something other than itself):
//   orc_philox4x32_10  - Random123 known-answer vectors and rocRAND's
//                        open-source host implementation (tests/test_oracle_rng.py)
//   orc_normals        - moments of 1e6 draws, exact 23-bit uniform grid,
//                        FP64 libm Box-Muller on the same uniforms, bit
//                        equality with the product's host build
//   orc_step_inside / orc_step_outside - closed-form ray/sphere hits on
//                        constructed rays, tangents, tol-band branches, and
//                        brute-force bisection (tests/test_oracle_primitives.py)
//   orc_walk           - invariants (confinement inside/outside, |x|<=L/2,
//                        displacement continuity across wraps), free diffusion
//                        MSD = 2Dt, restricted sphere/cylinder closed forms,
//                        T2 decay, candidate-mode equivalence, hand-traced
//                        single-collision cases (tests/test_oracle_walk.py)
//   orc_seed           - every accepted point lies in a target structure and
//                        in no earlier non-target one; uniformity
//   orc_build_reference_segments - brute-force recount of the reference's
//                        AABB rule (diffsim3d.py:107-174)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

// ---------------------------------------------------------------- constants
// sim3d_kernel.cu:34-35: `const float tol = 1e-4;` and `someLargeNumber =
// 1e6`, so the tolerance is the binary32 value nearest 1e-4
// (9.99999974737875e-05), widened exactly to double here.
constexpr double kTol = (double)1e-4f;
constexpr double kLarge = 1e6;
// sim3d_kernel.cu:373 (sub-step cap)
constexpr int kMaxIter = 100;

// Stream tags of the project's Philox counter (counter word 3).
constexpr uint32_t kStreamWalk = 0u;
constexpr uint32_t kStreamSeed = 1u;

// The reference's comparison macros (sim3d_kernel.cu:7-8), NaN behaviour kept.
inline double MINr(double a, double b) { return (a > b) ? b : a; }
inline double MAXr(double a, double b) { return (a < b) ? b : a; }

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Round multipliers and
// Weyl key increments are the published constants of the algorithm.
inline void philox_round(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k0;
  const uint32_t n2 = hi0 ^ c[3] ^ k1;
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// 23-bit uniform on (0,1): ((x >> 9) + 1/2) * 2^-23, exact in binary32.
inline float unif23(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.0f / 8388608.0f); }

// The walk's normals are a binary32 Box-Muller whose every operation is
// correctly rounded (this file is built with -ffp-contract=off), so the
// walker computes the same bits:
//   ln u: u = m 2^e, m folded into [1/sqrt2, sqrt2), s = (m-1)/(m+1),
//         ln u = e ln2 + 2 (s + s^3/3 + s^5/5 + s^7/7 + s^9/9)  (atanh series);
//   sin/cos(2 pi v): v = q/4 + a, |a| <= 1/8, Taylor series of sin(2 pi a)
//         to degree 9 and cos(2 pi a) to degree 10, then the quadrant q.
// The coefficients are the binary32 roundings of 1/(2k+1) and of
// (-1)^k (2 pi)^n / n!; ln 2 = 0.693145752 + 1.42860677e-6.
inline float ln_u32(float u) {
  int e = 0;
  float m = std::frexp(u, &e) * 2.0f;  // u = m 2^(e-1), m in [1, 2)
  e -= 1;
  if (m > 1.41421354f) {
    m *= 0.5f;
    e += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float z = s * s;
  float p = std::fma(z, 0.111111112f, 0.142857149f);
  p = std::fma(z, p, 0.200000003f);
  p = std::fma(z, p, 0.333333343f);
  p = std::fma(z, p, 1.0f);
  const float fe = (float)e;
  return std::fma(fe, 0.693145752f, std::fma(fe, 1.42860677e-06f, (2.0f * s) * p));
}

inline void sincos_2pi32(float v, float& sn, float& cs) {
  const float q = std::floor(v * 4.0f + 0.5f);
  const float a = v - q * 0.25f;
  const float z = a * a;
  float ps = std::fma(z, 42.0586929f, -76.7058563f);
  ps = std::fma(z, ps, 81.6052475f);
  ps = std::fma(z, ps, -41.3417015f);
  ps = std::fma(z, ps, 6.28318548f);
  const float sa = a * ps;
  float pc = std::fma(z, -26.4262562f, 60.2446404f);
  pc = std::fma(z, pc, -85.4568176f);
  pc = std::fma(z, pc, 64.9393921f);
  pc = std::fma(z, pc, -19.7392082f);
  const float ca = std::fma(z, pc, 1.0f);
  const int quad = ((int)q) & 3;
  sn = quad == 0 ? sa : quad == 1 ? ca : quad == 2 ? -sa : -ca;
  cs = quad == 0 ? ca : quad == 1 ? -sa : quad == 2 ? -ca : sa;
}

// Three standard normals for one (seed, stream, spin, step, draw) counter:
// Box-Muller on (u0,u1) gives n0,n1; on (u2,u3) gives n2.
inline void normals3(uint64_t seed, uint32_t stream, uint32_t spin, uint32_t step,
                     uint32_t draw, double out[3]) {
  uint32_t c[4] = {draw, step, spin, stream};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const float r0 = std::sqrt(-2.0f * ln_u32(unif23(c[0])));
  const float r1 = std::sqrt(-2.0f * ln_u32(unif23(c[2])));
  float s0, c0, s1, c1;
  sincos_2pi32(unif23(c[1]), s0, c0);
  sincos_2pi32(unif23(c[3]), s1, c1);
  out[0] = (double)(r0 * c0);
  out[1] = (double)(r0 * s0);
  out[2] = (double)(r1 * c1);
}

// ------------------------------------------------------ collision primitives
// isInsideSphere: sim3d_kernel.cu:72-75
inline double inside_sphere(double dx, double dy, double dz, double r) {
  return dx * dx + dy * dy + dz * dz - r * r;
}

// findStepSizeToSphereInsideTol: sim3d_kernel.cu:77-116
double step_inside(double dx, double dy, double dz, double r, double jx, double jy, double jz) {
  if ((dx * dx + dy * dy + dz * dz - r * r) > 0) return kLarge;       // :82-83
  const double a = jx * jx + jy * jy + jz * jz;                         // :86
  const double b = 2 * (jx * dx + jy * dy + jz * dz);                   // :87
  const double c = dx * dx + dy * dy + dz * dz - (r - kTol) * (r - kTol);  // :88
  double sq = b * b - 4 * a * c;                                        // :90
  if (sq < 0) return 0;                                                 // :96-97
  sq = std::sqrt(sq);                                                   // :101
  double q = -0.5 * (b + sq);                                           // :102
  if (b < 0) q = -0.5 * (b - sq);                                       // :103
  const double fpos = q / a, fneg = c / q;                              // :104-105
  return MAXr(MAXr(fpos, fneg), 0);                                     // :115
}

// findStepSizeToSphereOutsideTol: sim3d_kernel.cu:118-160
double step_outside(double dx, double dy, double dz, double r, double jx, double jy, double jz) {
  const double a = jx * jx + jy * jy + jz * jz;                         // :121
  const double b = 2 * (jx * dx + jy * dy + jz * dz);                   // :122
  const double c = dx * dx + dy * dy + dz * dz - (r + kTol) * (r + kTol);  // :123
  double sq = b * b - 4 * a * c;                                        // :125
  if (sq < 0) return kLarge;                                            // :129-130
  sq = std::sqrt(sq);                                                   // :135
  double q = -0.5 * (b + sq);                                           // :136
  if (b < 0) q = -0.5 * (b - sq);                                       // :137
  const double fpos = q / a, fneg = c / q;                              // :138-139
  if ((dx * dx + dy * dy + dz * dz - r * r) < 0) return MAXr(fpos, fneg);  // :141-145
  if (fpos > 0 && fneg > 0) return MINr(fpos, fneg);                    // :154-155
  if ((fpos < 0 && fneg < 0) || (b > 0)) return kLarge;                 // :156-157
  return 0;                                                             // :159
}

// ------------------------------------------------------------- substrate
struct Substrate {
  int S = 0;
  int nstruct = 0;
  std::vector<double> cx, cy, cz, r;
  std::vector<int> sid;
  std::vector<double> D, T2;  // nstruct + 1 entries, box last
  double L[3] = {0, 0, 0};

  // Candidate mode 0: exact bins (each sphere in every bin its tol-expanded
  // AABB touches) over a region around the box; segment queries gather the
  // bins under the segment's AABB.  Complete by construction; a query that
  // leaves the region falls back to all spheres.
  double blo[3], bsz;
  int bn[3];
  std::vector<int> bin_start, bin_items;

  // Candidate mode 1: the reference's padded per-segment lists
  // (diffsim3d.py:107-174), the cell chosen once per step from the pre-step
  // position (sim3d_kernel.cu:367-369).
  int mode = 0;
  int nseg[3] = {1, 1, 1};
  int sps = 0;
  std::vector<int> seglist;
};

void build_bins(Substrate& s, double bin_size, double halo) {
  s.bsz = bin_size;
  for (int a = 0; a < 3; ++a) {
    s.blo[a] = -s.L[a] / 2 - halo;
    s.bn[a] = std::max(1, (int)std::ceil((s.L[a] + 2 * halo) / bin_size));
  }
  const size_t nb = (size_t)s.bn[0] * s.bn[1] * s.bn[2];
  std::vector<int> cnt(nb + 1, 0);
  auto range = [&](int i, int a, int& lo, int& hi) {
    const double c = (a == 0 ? s.cx[i] : a == 1 ? s.cy[i] : s.cz[i]);
    const double ext = s.r[i] + 2 * kTol + 1e-9;
    lo = (int)std::floor((c - ext - s.blo[a]) / s.bsz);
    hi = (int)std::floor((c + ext - s.blo[a]) / s.bsz);
    lo = std::max(lo, 0);
    hi = std::min(hi, s.bn[a] - 1);
  };
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int> fill;
    if (pass == 1) {
      for (size_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
      s.bin_start = cnt;
      s.bin_items.assign(cnt[nb], 0);
      fill.assign(cnt.begin(), cnt.end() - 1);
    }
    for (int i = 0; i < s.S; ++i) {
      int lo[3], hi[3];
      bool ok = true;
      for (int a = 0; a < 3; ++a) {
        range(i, a, lo[a], hi[a]);
        if (lo[a] > hi[a]) ok = false;
      }
      if (!ok) continue;
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) {
            const size_t b = ((size_t)z * s.bn[1] + y) * s.bn[0] + x;
            if (pass == 0) cnt[b + 1]++;
            else s.bin_items[fill[b]++] = i;
          }
    }
  }
}

// Gather candidates for the segment p -> e (mode 0).
void query_exact(const Substrate& s, const double p[3], const double e[3], std::vector<int>& out) {
  out.clear();
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    const double mn = std::min(p[a], e[a]) - 2 * kTol, mx = std::max(p[a], e[a]) + 2 * kTol;
    const double flo = std::floor((mn - s.blo[a]) / s.bsz), fhi = std::floor((mx - s.blo[a]) / s.bsz);
    if (!(flo >= 0 && fhi < s.bn[a])) {  // leaves the binned region: every sphere
      out.resize(s.S);
      for (int i = 0; i < s.S; ++i) out[i] = i;
      return;
    }
    lo[a] = (int)flo;
    hi[a] = (int)fhi;
  }
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const size_t b = ((size_t)z * s.bn[1] + y) * s.bn[0] + x;
        for (int k = s.bin_start[b]; k < s.bin_start[b + 1]; ++k) out.push_back(s.bin_items[k]);
      }
}

// Reference segment index from the pre-step position (sim3d_kernel.cu:367-369).
int reference_segment(const Substrate& s, const double p[3]) {
  int idx[3];
  for (int a = 0; a < 3; ++a) {
    const double dL = s.L[a] / s.nseg[a];
    const double f = std::floor((p[a] + s.L[a] / 2) / dL);
    idx[a] = (int)MINr((double)(s.nseg[a] - 1), MAXr(0.0, f));
  }
  return idx[2] * s.nseg[0] * s.nseg[1] + idx[1] * s.nseg[0] + idx[0];
}

struct PhaseIn {
  const double* g = nullptr;      // per-step phase coefficient (diffsim3d.py:254)
  const double* garea = nullptr;  // per-step wrap area (diffsim3d.py:255)
};

struct SpinOut {
  long long nsub = 0;       // sub-steps beyond the first (scatter events)
  long long ncap = 0;       // steps that hit the 100-iteration cap
};

// One spin through nsteps (sim3d_kernel.cu:340-522 with the phase additions of
// :719-943).  pos/pos0 are absolute, wrapped exactly as the reference wraps.
void walk_spin(const Substrate& s, uint64_t seed, uint32_t spin, long long step0, int nsteps,
               double dt, double* pos, double* pos0, double* sig, double* mom, const PhaseIn& ph,
               double* stepmom /* nsteps x 6, per-spin contributions, may be null */,
               SpinOut& so, std::vector<int>& cand) {
  double p[3] = {pos[0], pos[1], pos[2]};
  double p0[3] = {pos0[0], pos0[1], pos0[2]};
  double sg = *sig;
  for (int t = 0; t < nsteps; ++t) {
    const uint32_t step = (uint32_t)(step0 + t);
    double ds = 0, j[3] = {0, 0, 0};
    double jr = 1.0;   // :364
    int niter = 0;
    // mode 1: cell from the pre-step position, fixed for the step (:367-369)
    const int seg = (s.mode == 1) ? reference_segment(s, p) : 0;
    const int* segl = (s.mode == 1) ? &s.seglist[(size_t)seg * s.sps] : nullptr;
    while (jr > kTol && niter < kMaxIter) {  // :373
      niter++;
      double jc = jr;  // :376
      // Candidate set for this sub-step.  Mode 0 needs the segment; the jump
      // of the first sub-step is only known after the structure pass, whose
      // candidates are those containing p, so gather for the point first and
      // for the full segment below.
      int sid = s.nstruct;  // :379
      auto scan_sid = [&](const int* list, int n) {
        for (int k = 0; k < n; ++k) {
          const int i = list[k];
          if (s.sid[i] < sid) {                                                // :384
            if (inside_sphere(p[0] - s.cx[i], p[1] - s.cy[i], p[2] - s.cz[i], s.r[i]) < 0)  // :387-393
              sid = s.sid[i];                                                 // :395
          }
        }
      };
      if (s.mode == 1) {
        scan_sid(segl, s.sps);
      } else {
        query_exact(s, p, p, cand);
        scan_sid(cand.data(), (int)cand.size());
      }
      if (niter == 1) {  // :400-406
        ds = std::sqrt(2 * s.D[sid] * dt);
        double n[3];
        normals3(seed, kStreamWalk, spin, step, 0u, n);
        j[0] = ds * n[0]; j[1] = ds * n[1]; j[2] = ds * n[2];
      }
      double insideAny = 0, fin = 0, fout = kLarge;  // :408-411
      auto scan_col = [&](const int* list, int n) {
        for (int k = 0; k < n; ++k) {
          const int i = list[k];
          if (s.sid[i] > sid) continue;                                         // :418-419
          const double dx = p[0] - s.cx[i], dy = p[1] - s.cy[i], dz = p[2] - s.cz[i];
          const double isIn = inside_sphere(dx, dy, dz, s.r[i]);               // :422-426
          if (isIn < 0) {                                                       // :428
            insideAny = 1;
            const double f = step_inside(dx, dy, dz, s.r[i], j[0], j[1], j[2]);
            if (fin < f) fin = f;                                               // :438-439
          } else if (s.sid[i] < sid) {                                          // :441
            const double f = step_outside(dx, dy, dz, s.r[i], j[0], j[1], j[2]);
            if (f < fout) fout = f;                                             // :452-453
          }
        }
      };
      if (s.mode == 1) {
        scan_col(segl, s.sps);
      } else {
        const double e[3] = {p[0] + jr * j[0], p[1] + jr * j[1], p[2] + jr * j[2]};
        query_exact(s, p, e, cand);
        scan_col(cand.data(), (int)cand.size());
      }
      jc = MINr(jc, insideAny ? fin : kLarge);  // :457
      jc = MINr(jc, fout);                      // :458
      p[0] += jc * j[0]; p[1] += jc * j[1]; p[2] += jc * j[2];  // :460-462
      if (jc < jr) {  // :464-479 scatter
        jr = std::sqrt(MAXr(0, jr * jr - jc * jc));
        sg *= std::exp(-jc * jc * dt / s.T2[sid]);
        double n[3];
        normals3(seed, kStreamWalk, spin, step, (uint32_t)niter, n);
        j[0] = ds * n[0]; j[1] = ds * n[1]; j[2] = ds * n[2];
        so.nsub++;
      } else {  // :480-486 coherent
        jr -= jc;
        sg *= std::exp(-jc * dt / s.T2[sid]);
      }
    }
    if (niter >= kMaxIter && jr > kTol) so.ncap++;
    // phase after the move, before the wrap (:880-887)
    if (ph.g && mom) {
      const double G = ph.g[t];
      mom[0] += p[0] * G; mom[1] += p[1] * G; mom[2] += p[2] * G;
    }
    // periodic wrap, once per axis, with the spins0 shift (:489-521) and the
    // phase correction (:891-942)
    for (int a = 0; a < 3; ++a) {
      const double La = s.L[a];
      if (p[a] < -La / 2) {
        p[a] += La; p0[a] += La;
        if (ph.g && mom) mom[a] += La * ph.garea[t];
      } else if (p[a] > La / 2) {
        p[a] -= La; p0[a] -= La;
        if (ph.g && mom) mom[a] -= La * ph.garea[t];
      }
    }
    if (stepmom) {  // computeDiffusionCoefficientsAndKurtosis: sim3d_kernel.cu:191-200
      for (int a = 0; a < 3; ++a) {
        const double d = p[a] - p0[a];
        stepmom[t * 6 + a] = d * d;
        stepmom[t * 6 + 3 + a] = d * d * d * d;
      }
    }
  }
  pos[0] = p[0]; pos[1] = p[1]; pos[2] = p[2];
  pos0[0] = p0[0]; pos0[1] = p0[1]; pos0[2] = p0[2];
  *sig = sg;
}

struct Handle {
  Substrate s;
};

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// n draws of three normals: draw k uses counter (draw0 + k, step, spin, stream).
void orc_normals(uint64_t seed, uint32_t stream, uint32_t spin, uint32_t step, uint32_t draw0,
                 int n, double* out) {
  for (int k = 0; k < n; ++k) normals3(seed, stream, spin, step, draw0 + (uint32_t)k, out + 3 * k);
}

double orc_step_inside(double dx, double dy, double dz, double r, double jx, double jy, double jz) {
  return step_inside(dx, dy, dz, r, jx, jy, jz);
}
double orc_step_outside(double dx, double dy, double dz, double r, double jx, double jy, double jz) {
  return step_outside(dx, dy, dz, r, jx, jy, jz);
}

// spheres: [5, S] row-major (cx, cy, cz, r, sid) as SimGeometry3D.spheres
// (geometry.py:43,58-64).  D, T2: nstruct + 1 entries with the box last
// (diffsim3d.py:95-103).
void* orc_substrate_create(const double* spheres, int S, int nstruct, const double* D,
                           const double* T2, double Lx, double Ly, double Lz, double bin_size,
                           double halo) {
  Handle* h = new Handle();
  Substrate& s = h->s;
  s.S = S;
  s.nstruct = nstruct;
  s.L[0] = Lx; s.L[1] = Ly; s.L[2] = Lz;
  s.cx.assign(spheres, spheres + S);
  s.cy.assign(spheres + S, spheres + 2 * S);
  s.cz.assign(spheres + 2 * S, spheres + 3 * S);
  s.r.assign(spheres + 3 * S, spheres + 4 * S);
  s.sid.resize(S);
  for (int i = 0; i < S; ++i) s.sid[i] = (int)spheres[4 * S + i];
  s.D.assign(D, D + nstruct + 1);
  s.T2.assign(T2, T2 + nstruct + 1);
  build_bins(s, bin_size > 0 ? bin_size : 1.0, halo > 0 ? halo : 4.0);
  return h;
}

void orc_substrate_destroy(void* h) { delete (Handle*)h; }

// Switch to the reference's padded segment lists (mode 1) or back to the
// exact bins (mode 0).  seglist: [nx*ny*nz, sps] int32 as DiffSim3d.segments.
void orc_substrate_set_mode(void* hh, int mode, int nx, int ny, int nz, int sps, const int* seglist) {
  Substrate& s = ((Handle*)hh)->s;
  s.mode = mode;
  if (mode == 1) {
    s.nseg[0] = nx; s.nseg[1] = ny; s.nseg[2] = nz;
    s.sps = sps;
    s.seglist.assign(seglist, seglist + (size_t)nx * ny * nz * sps);
  }
}

// The walker over nsteps for N spins (OpenMP over spins).
//   pos, pos0 : [3, N] in/out (absolute, wrapped), sig : [N] in/out
//   g, garea  : [nsteps] or null; mom : [3, N] in/out or null
//   stepmom   : [nsteps, 6] out or null (sums over spins of dx^2,dy^2,dz^2,dx^4,dy^4,dz^4)
//   nsub      : [N] out or null (scatter sub-steps per spin, accumulated)
//   stats     : [2] out or null (total sub-steps, capped steps)
int orc_walk(void* hh, int N, double* pos, double* pos0, double* sig, uint64_t seed,
             long long step0, int nsteps, double dt, const double* g, const double* garea,
             double* mom, double* stepmom, long long* nsub, long long* stats, int nthreads) {
  const Substrate& s = ((Handle*)hh)->s;
  PhaseIn ph;
  ph.g = g;
  ph.garea = garea;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  long long tot_sub = 0, tot_cap = 0;
  const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
  std::vector<double> smom_thr(stepmom ? (size_t)nth * nsteps * 6 : 0, 0.0);
#pragma omp parallel reduction(+ : tot_sub, tot_cap)
  {
    std::vector<int> cand;
    std::vector<double> sm(stepmom ? (size_t)nsteps * 6 : 0);
    double* mythr = stepmom ? &smom_thr[(size_t)omp_get_thread_num() * nsteps * 6] : nullptr;
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
      double p[3] = {pos[i], pos[N + i], pos[2 * N + i]};
      double p0[3] = {pos0[i], pos0[N + i], pos0[2 * N + i]};
      double m[3] = {0, 0, 0};
      if (mom) { m[0] = mom[i]; m[1] = mom[N + i]; m[2] = mom[2 * N + i]; }
      SpinOut so;
      walk_spin(s, seed, (uint32_t)i, step0, nsteps, dt, p, p0, &sig[i], mom ? m : nullptr, ph,
                stepmom ? sm.data() : nullptr, so, cand);
      pos[i] = p[0]; pos[N + i] = p[1]; pos[2 * N + i] = p[2];
      pos0[i] = p0[0]; pos0[N + i] = p0[1]; pos0[2 * N + i] = p0[2];
      if (mom) { mom[i] = m[0]; mom[N + i] = m[1]; mom[2 * N + i] = m[2]; }
      if (nsub) nsub[i] += so.nsub;
      tot_sub += so.nsub;
      tot_cap += so.ncap;
      if (stepmom)
        for (size_t k = 0; k < sm.size(); ++k) mythr[k] += sm[k];
    }
  }
  if (stepmom) {
    for (size_t k = 0; k < (size_t)nsteps * 6; ++k) {
      double acc = 0;
      for (int t = 0; t < nth; ++t) acc += smom_thr[(size_t)t * nsteps * 6 + k];
      stepmom[k] = acc;
    }
  }
  if (stats) { stats[0] = tot_sub; stats[1] = tot_cap; }
  return 0;
}

// Rejection seeding (seedSpinsKernel, sim3d_kernel.cu:240-337, driven by
// geometry.py:137-168): a uniform point in the box is accepted when it lies
// in a target structure and in no earlier non-target structure.  Attempt k of
// spin i draws from counter (k, 0, i, stream 1).  Unlike the reference, the
// sphere index is advanced correctly after a hit (reference quirk at :299-305).
//   targets: ntarget structure ids (box = nstruct).  pos: [3, N] out.
//   Returns the number of spins still unseeded after max_attempts.
int orc_seed(void* hh, int N, const int* targets, int ntarget, uint64_t seed, int max_attempts,
             double* pos, int nthreads) {
  const Substrate& s = ((Handle*)hh)->s;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int failed = 0;
#pragma omp parallel reduction(+ : failed)
  {
    std::vector<char> inside(s.nstruct + 1);
    std::vector<int> cand;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
      bool done = false;
      for (int k = 0; k < max_attempts && !done; ++k) {
        uint32_t c[4] = {(uint32_t)k, 0u, (uint32_t)i, kStreamSeed};
        philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        double p[3];
        for (int a = 0; a < 3; ++a) p[a] = (unif23(c[a]) - 0.5) * s.L[a];
        std::fill(inside.begin(), inside.end(), 0);
        query_exact(s, p, p, cand);
        for (int idx : cand)
          if (inside_sphere(p[0] - s.cx[idx], p[1] - s.cy[idx], p[2] - s.cz[idx], s.r[idx]) < 0)
            inside[s.sid[idx]] = 1;
        // the box: strict inequalities (geometry.py:69-71, sim3d_kernel.cu:282-286)
        inside[s.nstruct] = (p[0] > -s.L[0] / 2 && p[0] < s.L[0] / 2 && p[1] > -s.L[1] / 2 &&
                             p[1] < s.L[1] / 2 && p[2] > -s.L[2] / 2 && p[2] < s.L[2] / 2);
        bool excluded = false, seeded = false;
        for (int st = 0; st <= s.nstruct; ++st) {
          bool target = false;
          for (int q = 0; q < ntarget; ++q) target |= (targets[q] == st);
          if (target) {
            if (inside[st] && !excluded) seeded = true;
          } else {
            excluded = excluded || inside[st];
          }
        }
        if (seeded) {
          pos[i] = p[0]; pos[N + i] = p[1]; pos[2 * N + i] = p[2];
          done = true;
        }
      }
      if (!done) failed++;
    }
  }
  return failed;
}

}  // extern "C"
