/*
 * nozomi_hip.h - C ABI of libnozomi_hip.so, the MI355X (gfx950) diffusion-MRI
 * random walker.
 *
 * The reference drives its walker through PyCUDA: it %-formats the CUDA
 * template simulation_toolkit/simulation_engine/sim3d_kernel.cu with the spin
 * count, sphere count, box and per-structure D/T2 (diffsim3d.py:35-63,84-104),
 * JIT-compiles it and launches one kernel per time step through
 * get_function(...) handles.  This library replaces that "operator layer"
 * with ordinary entry points: run-time parameters instead of baked-in
 * constants, many steps per launch, and caller-owned device buffers.
 *
 * Conventions
 *  - Every float / int pointer argument is a DEVICE pointer (for example
 *    torch.Tensor.data_ptr() of a tensor on cuda:0) unless the argument
 *    name ends in _host.
 *  - Arrays are structure-of-arrays in the reference's layouts: [3, N]
 *    positions are x[0..N), y[0..N), z[0..N) back to back.
 *  - stream is a hipStream_t (NULL = the null stream).  Calls are asynchronous
 *    on that stream except where stated; nothing synchronises implicitly.
 *  - Return value: 0 on success, < 0 for an argument / shape / budget error,
 *    > 0 for a hipError_t.  nz_last_error_string() gives the message
 *    (thread-local).
 *  - The library never frees caller memory.  It owns only the opaque
 *    substrate handle (device copy of the spheres plus the cell grid).
 */
#ifndef NOZOMI_HIP_H
#define NOZOMI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NZ_ABI_VERSION 5

typedef struct nz_substrate* nz_substrate_t;

/* Grid and substrate figures, filled by nz_substrate_info. */
typedef struct nz_substrate_info_t {
  int64_t nspheres;
  int32_t nstructures;
  int32_t ncell[3];        /* cells per axis inside the box              */
  double cell[3];          /* cell edge per axis (um)                    */
  double margin;           /* safety margin of the per-cell lists (um)   */
  int64_t ncand_total;     /* candidate entries over all cells           */
  int32_t ncand_max;       /* largest list                               */
  double ncand_mean_box;   /* mean list length over the in-box cells     */
  int64_t device_bytes;    /* device memory held by the handle           */
  int64_t ncand_records;   /* records in the grid's record array, the
                              padding records after the last list included */
} nz_substrate_info_t;

int nz_abi_version(void);
const char* nz_last_error_string(void);

/* Upload a substrate and build its cell grid.
 * Replaces SimGeometry3D.spheres upload + DiffSim3d.set_segments
 * (diffsim3d.py:54,107-174, geometry.py:37-67).
 *   spheres_host : [5, nspheres] row-major doubles (cx, cy, cz, r, sid) as
 *                  SimGeometry3D.spheres; sid = structure index (precedence:
 *                  lower wins, sim3d_kernel.cu:53-54)
 *   D_host, T2_host : nstructures + 1 entries, the box last (diffsim3d.py:95-103)
 *   L*       : box edges (um); periodic cube [-L/2, L/2]
 *   cell_size: target cell edge (um, 0 = 2.0); the grid tiles the box exactly
 *   margin   : per-cell list margin (um, 0 = 0.8); a sub-step that would
 *              leave the cell's box grown by the margin re-bins the spin
 * Synchronous: validates on the host, builds the lists, records and sort
 * ranks on the current device (the reference builds its lists on the host). */
int nz_substrate_create(const double* spheres_host, int64_t nspheres, int32_t nstructures,
                        const double* D_host, const double* T2_host, double Lx, double Ly,
                        double Lz, double cell_size, double margin, nz_substrate_t* out);
int nz_substrate_destroy(nz_substrate_t sub);
int nz_substrate_info(nz_substrate_t sub, nz_substrate_info_t* info);

/* Copy a substrate's cell grid to host memory (inspection and tests; no
 * counterpart in the reference, whose segment table is a host array,
 * diffsim3d.py:166-174).  Each pointer may be NULL; sizes from
 * nz_substrate_info with nloc = ncell + 2 per axis:
 *   cell_start : int32 [nloc_x nloc_y nloc_z + 1], CSR list starts (cells
 *                indexed (z nloc_y + y) nloc_x + x, halo cells included)
 *   cell_rank  : int32 [ncell_x ncell_y ncell_z], sort rank of each in-box cell
 *   cand       : float [8 ncand_records], records (cx, cy, cz, (r+tol)^2)
 *                and (r^2, (r-tol)^2, sid, sphere index bits) per entry,
 *                coordinates relative to the cell's lower corner, then
 *                ncand_records - ncand_total zero padding records.
 *                Synchronous. */
int nz_substrate_export_grid(nz_substrate_t sub, int32_t* cell_start, int32_t* cell_rank,
                             float* cand);

/* Copy a substrate's reach entries to host memory (inspection and tests; no
 * counterpart in the reference).  The candidate scan's first pass reads one
 * 8-byte entry per list entry: binary16 (h_x, h_y, h_z, k) with h = -2 sigma
 * c' (c' the centre relative to the cell's lower corner as h encodes it) and
 * k at most sigma (|c'|^2 - (1 + lambda) R'^2), R' = r + tol + |c - c'|.
 * Each pointer may be NULL:
 *   nentries    : int64, the number of entries (lists padded to even
 *                 lengths, then 16 zero entries)
 *   reach_start : int32 [nloc_x nloc_y nloc_z], first entry of each cell's list
 *   reach       : uint32 [2 nentries], h_x | h_y << 16 and h_z | k << 16.
 * Synchronous. */
int nz_substrate_export_reach(nz_substrate_t sub, int64_t* nentries, int32_t* reach_start,
                              uint32_t* reach);

/* Rejection-seed n spins uniformly inside the target structures.
 * Replaces SimGeometry3D.gpu_seed + seedSpinsKernel (geometry.py:73-174,
 * sim3d_kernel.cu:240-337).  A point is accepted when it lies in a target
 * structure and in no earlier non-target structure.  Attempt k of spin i
 * draws Philox counter (k, 0, i, 1) under key seed.
 *   targets_host : ntargets structure ids (box = nstructures)
 *   pos          : [3, n] device output
 *   nfailed      : device int32, receives the number of spins still unplaced
 *                  after max_attempts (may be NULL) */
int nz_seed(nz_substrate_t sub, int64_t n, const int32_t* targets_host, int32_t ntargets,
            uint64_t seed, int32_t max_attempts, float* pos, int32_t* nfailed, void* stream);

/* A walker: the resident state of n spins on the substrate's device.
 * Replaces the device arrays DiffSim3d.setup allocates (spins_d, spins0_d,
 * sig_d, phase_d and the cuRAND states; diffsim3d.py:35-63).  The state is
 * kept in cell-relative double precision, ordered by cell (re-sorted every
 * sort_interval steps so that a wavefront's spins share candidate lists);
 * spin i keeps its Philox key i whatever its slot.  Results come back in spin
 * order through nz_walker_read.  Library-owned; device memory ~210 B/spin. */
typedef struct nz_walker* nz_walker_t;

#define NZ_WALKER_COUNTERS 1u  /* keep per-spin scatter and FP64-rescan counters */

int nz_walker_create(nz_substrate_t sub, int64_t n, uint32_t flags, nz_walker_t* out);
int nz_walker_destroy(nz_walker_t w);
/* Steps between spatial re-sorts of the spins (0 = never sort; default 256). */
int nz_walker_set_sort_interval(nz_walker_t w, int32_t steps);
/* Global index of the next step (the Philox step counter); set it to resume. */
int nz_walker_get_step(nz_walker_t w, int64_t* step);
int nz_walker_set_step(nz_walker_t w, int64_t step);

/* Start positions pos [3,n] (absolute, device): resets the state (pos0 = pos,
 * sig = 1, phase moment 0, counters 0, step 0) and finds every spin's
 * structure id.  Replaces the set-up of DiffSim3d.setup (diffsim3d.py:42-62);
 * initstates (sim3d_kernel.cu:65-70) has no counterpart: the Philox stream is
 * stateless. */
int nz_walker_set_positions(nz_walker_t w, const float* pos, void* stream);

/* Rejection-seed the walker's spins into the target structures (as nz_seed)
 * and set them as start positions.  nfailed: device int32 (may be NULL). */
int nz_walker_seed(nz_walker_t w, const int32_t* targets_host, int32_t ntargets, uint64_t seed,
                   int32_t max_attempts, int32_t* nfailed, void* stream);

/* Advance every spin by nsteps time steps of length dt (ms).
 * Replaces nsteps launches of randomWalk3d / randomWalk3d_phase /
 * randomWalk3d_phase_multidirr (sim3d_kernel.cu:340-943, launched by
 * diffsim3d.py:199-226) with persistent launches of many steps each.
 *   seed   : Philox key; draw k of step s of spin i uses counter (k, s, i, 0),
 *            s counting from the walker's step counter, which advances
 *   g, garea : device float [nsteps] per-step phase coefficient and wrap area,
 *              i.e. the G and G_area arguments of the reference launches
 *              (diffsim3d.py:253-255); NULL for a walk without phase.
 *              With them the walker's moment accumulates sum_t g_t x_t with
 *              the reference's wrap correction (sim3d_kernel.cu:674-715,880-942).
 *   stepmom: device double [nsteps, 6] or NULL; receives (adds) the sums over
 *            spins of dx^2, dy^2, dz^2, dx^4, dy^4, dz^4 after every step (the
 *            moment sums of computeDiffusionCoefficientsAndKurtosis,
 *            sim3d_kernel.cu:162-238).
 *   flags  : NZ_WALK_* bits. */
#define NZ_WALK_T2 1u          /* update the T2 weight (randomWalk3d's sig) */
int nz_walk(nz_walker_t w, uint64_t seed, int32_t nsteps, double dt, const float* g,
            const float* garea, double* stepmom, uint32_t flags, void* stream);

/* The reference's views in spin order, any pointer may be NULL:
 * pos [3,n] wrapped positions (spins_d), pos0 [3,n] start positions shifted by
 * every wrap (spins0_d), sig [n] T2 weight (sig_d), mom [3,n] phase moment. */
int nz_walker_read(nz_walker_t w, float* pos, float* pos0, float* sig, float* mom, void* stream);
/* FP64 wrapped positions and unwrapped displacements [3,n] in spin order. */
int nz_walker_read_f64(nz_walker_t w, double* pos, double* disp, void* stream);
/* Per-spin counters [n] (walker created with NZ_WALKER_COUNTERS): scatter
 * sub-steps and sub-steps whose candidate scan ran in FP64. */
int nz_walker_counters(nz_walker_t w, int32_t* nsub, int32_t* nguard, void* stream);

/* Sums over spins of dx^2, dy^2, dz^2, dx^4, dy^4, dz^4 of the current
 * displacement (position minus spins0): the block sums of
 * computeDiffusionCoefficientsAndKurtosis (sim3d_kernel.cu:162-238) before the
 * host scaling D = S/(2tN), K = S/N.  out: device double [6], overwritten. */
int nz_displacement_moments(nz_walker_t w, double* out, void* stream);

/* Launch timing of the walk kernel (no counterpart in the reference, which
 * has no timing inside its launches): with enable != 0 every walk-kernel
 * launch of nz_walk is bracketed by HIP events on the launch stream.
 * nz_walker_kernel_time waits for the recorded launches, returns their total
 * duration (ms) and count, and starts a new tally. */
int nz_walker_time_kernels(nz_walker_t w, int32_t enable);
int nz_walker_kernel_time(nz_walker_t w, double* ms_total, int64_t* launches);

/* ---- many simulations in one launch (no counterpart in the reference) ----
 * The batched and served forms of the walker: a pool of spin slots shared by
 * up to max_sims independent simulations, each with its own substrate, spin
 * count, Philox key, time step, length and phase coefficients.  One launch
 * advances every active simulation by up to its k steps; a spin's trajectory
 * is the one nz_walk gives for the same substrate, key and start position
 * (counter-based RNG), whatever else shares the launch.  The sort/walk/advance
 * sequence runs on the batch's own stream, captured in a HIP graph, ordered
 * after and before the caller's stream by events. */
typedef struct nz_batch* nz_batch_t;
#define NZ_EFULL (-2)   /* nz_batch_add: no free simulation slot or spin blocks */

/* capacity: spin slots (rounded up to 256); max_sims: concurrent simulations. */
int nz_batch_create(int64_t capacity, int32_t max_sims, nz_batch_t* out);
int nz_batch_destroy(nz_batch_t b);
/* Add a simulation of n spins on sub: nsteps steps of dt (ms) with Philox key
 * seed, optional per-step phase coefficients g/garea (device float [nsteps],
 * kept by the caller until the simulation is removed), NZ_WALK_* flags, k
 * steps per launch.  Start positions: pos (device [3,n]) or, when pos is NULL,
 * rejection seeding into targets_host with key seed ^ 0x5EED5EED5EED (as
 * DiffSim3d.setup).  Returns NZ_EFULL when the pool has no room. */
int nz_batch_add(nz_batch_t b, nz_substrate_t sub, int64_t n, uint64_t seed, int64_t nsteps,
                 double dt, const float* g, const float* garea, uint32_t flags, int32_t k,
                 const float* pos, const int32_t* targets_host, int32_t ntargets, int32_t* sim,
                 void* stream);
int nz_batch_remove(nz_batch_t b, int32_t sim);
/* Enqueue `launches` launches (pairs replay the captured graph). */
int nz_batch_run(nz_batch_t b, int32_t launches, void* stream);
int nz_batch_set_graph(nz_batch_t b, int32_t enable);
/* Steps completed and whether the simulation finished (waits for the batch). */
int nz_batch_status(nz_batch_t b, int32_t sim, int64_t* step, int32_t* done);
/* The same for every simulation entry at once (arrays of max_sims; unused
 * entries report step -1, done -1). */
int nz_batch_status_all(nz_batch_t b, int64_t* steps, int32_t* done);
/* Free capacity: largest contiguous run of free slots, all free slots, free
 * simulation entries. */
int nz_batch_free(nz_batch_t b, int64_t* largest_run, int64_t* free_slots, int32_t* free_sims);
/* A simulation's state in spin order (as nz_walker_read). */
int nz_batch_read(nz_batch_t b, int32_t sim, float* pos, float* pos0, float* sig, float* mom,
                  void* stream);
/* A simulation's DWI signal sums (as nz_signal_reduce over its moments). */
int nz_batch_signals(nz_batch_t b, int32_t sim, int32_t W, const float* dirs, const float* scale,
                     double* out, void* stream);

/* ---- precision form: blocked phase contraction of many waveforms ----------
 * (no counterpart in the reference, whose phase kernels take one waveform
 * shape per launch, sim3d_kernel.cu:719-943).  For W waveform shapes g_w(t)
 * with directions u_w: phase[i,w] = (x0_i . u_w) sum_t g_w(t) (nz_phase_init)
 * + sum over K-step blocks of rec . Bt (nz_phase_gemm), where rec holds the
 * block's unwrapped displacements (nz_walk_record) and Bt[w][3(t-t0)+a] =
 * g_w(t) u_w,a (nz_waveform_block). */
/* Walk nsteps (<= 4096) with every spin stopping at the block end, recording
 * FP32 displacements from the start: rec [n][ld] slot-major, entry 3t + a of
 * slot s the displacement along axis a after step t of the block (ld >= 3
 * nsteps floats per slot); slot_spin [n] receives the slot -> spin map of the
 * record. */
int nz_walk_record(nz_walker_t w, uint64_t seed, int32_t nsteps, double dt, float* rec,
                   int64_t ld, int32_t* slot_spin, void* stream);
/* nz_walk_record into a ring of `ring` steps per slot (ld >= 3 ring): step
 * t of the block goes to entries 3 e + a, e = (t - t0 + pos0) mod ring, the
 * block itself to entries pos0 .. pos0 + nsteps - 1 (pos0 + nsteps <= ring),
 * so a contraction reads it at rec + 3 pos0 with the same ld.  A spin whose
 * block is done walks on, up to ahead further steps (nsteps + ahead <= ring),
 * while the other spins of its wavefront finish theirs, recording those steps
 * at the following ring entries; the tail of the wavefront's block then walks
 * the next block instead of idling.  The next record call continues in the
 * same ring at entry (pos0 + nsteps) mod ring and is at least ahead steps
 * long.  No re-sort of the spins runs while steps run ahead sit in the ring:
 * the block before a due re-sort runs no steps ahead.  The caller keeps the
 * entries a contraction still reads out of the next ahead entries (with the
 * contraction of block b beside the walk of block b + 1: ring >= 3 blocks).
 * A plain nz_walk, or new positions or steps, end the run-ahead. */
int nz_walk_record_ring(nz_walker_t w, uint64_t seed, int32_t nsteps, int32_t ahead, double dt,
                        float* rec, int64_t ld, int32_t ring, int32_t pos0, int32_t* slot_spin,
                        void* stream);
/* Waveform block for steps t0..t0+K-1: g [W][T] per-step coefficients,
 * u [3][W] directions; Bt as BF16 hi/lo [W][3K] (uint16 bits) and, if
 * bt_f32 is not NULL, FP32. */
int nz_waveform_block(const float* g, int64_t T, const float* u, int32_t W, int64_t t0, int32_t K,
                      uint16_t* bt_hi, uint16_t* bt_lo, float* bt_f32, void* stream);
/* phase [n_spins][W] += rec . Bt^T over the block, rec [nslots][ld] as
 * nz_walk_record writes it (ld >= 3K): mode 0 = split hi/lo BF16 MFMA (K =
 * 32, 64, 96 or 128; W and ld multiples of 4), mode 1 = FP32 vector ALUs. */
int nz_phase_gemm(const float* rec, int64_t ld, int64_t nslots, const int32_t* slot_spin,
                  int32_t K, const uint16_t* bt_hi, const uint16_t* bt_lo, const float* bt_f32,
                  int32_t W, int32_t mode, float* phase, void* stream);
/* phase [n][W] = (x0_i . u_w) gsum[w]; x0 [3][n] start positions. */
int nz_phase_init(const float* x0, int64_t n, const float* u, const float* gsum, int32_t W,
                  float* phase, void* stream);
/* out [2W] = sum_i cos / sin phase[i][w], FP64. */
int nz_phase_signals(const float* phase, int64_t n, int32_t W, double* out, void* stream);

/* Phases of W directions from the moments: phase[w, i] = scale[w] *
 * (dirs[0,w] m_x + dirs[1,w] m_y + dirs[2,w] m_z).  Replaces the per-step
 * W-wide accumulation of randomWalk3d_phase_multidirr (sim3d_kernel.cu:880-942);
 * the result equals DwiSim3d.simulate_multi_directions (diffsim3d.py:246-261).
 *   mom [3,n], dirs [3,W], scale [W] (NULL = 1), phase out [W, n]. */
int nz_project_phase(int64_t n, int32_t W, const float* mom, const float* dirs,
                     const float* scale, float* phase, void* stream);

/* DWI signal: out[2w] = sum_i cos(phase[w,i]), out[2w+1] = sum_i sin(phase[w,i])
 * with phase as in nz_project_phase, accumulated in FP64.  The reference has no
 * counterpart (its caller is missing); this is the signal step that turns
 * DwiSim3d phases into per-direction signals.  out: device double [2W],
 * overwritten. */
int nz_signal_reduce(int64_t n, int32_t W, const float* mom, const float* dirs,
                     const float* scale, double* out, void* stream);

/* The same signal reduction without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-block FP64 partials live in a static device array of
 * 1024 x 2 x 512 doubles (8 MiB), one copy per GPU, held until the process
 * ends; W is limited to 512 (other sizes return < 0 with a message).  The
 * result equals nz_signal_reduce's bit for bit.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller. */
int nz_signal_reduce_noalloc(int64_t n, int32_t W, const float* mom, const float* dirs,
                             const float* scale, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NOZOMI_HIP_H */
