"""Benchmark: spin-steps/s of the DWI walker on one MI355X.

Workload (one "step" of this benchmark = one full simulation pass):
200 beaded axons (seeded synthetic substrate, ~7e4 spheres incl. images) in
a 60 um periodic cube; 1e6 spins uniform in the box (intra/extra split by
position), D = 2.0 um^2/ms everywhere; dt = 10 us, 5000 time steps; PGSE
little_delta = 10 ms, big_delta = 40 ms at b = 1, 2, 3 ms/um^2 with 64
Fibonacci-hemisphere directions per shell (192 signals).  Inputs are resident
in HBM before timing; each timed pass re-initialises the spin state from the
same start positions (nz_walker_set_positions), walks 5000 steps with phase
(persistent launches between spatial re-sorts) and reduces the 192 signals in
FP64.

Prints one JSON line (rank 0).  The CPU baseline is the FP64 oracle (OpenMP)
on a bounded subsample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3
VALU_ISSUE_PER_S = 256 * 2.4e9  # wave64 VALU instructions: one per CU per cycle at 2.4 GHz


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spins", type=int, default=1_000_000)
    ap.add_argument("--nsteps", type=int, default=5000, help="time steps per simulation")
    ap.add_argument("--axons", type=int, default=200)
    ap.add_argument("--box", type=float, default=60.0)
    ap.add_argument("--cell", type=float, default=0.0)
    ap.add_argument("--margin", type=float, default=0.0)
    ap.add_argument("--sort", type=int, default=250,
                    help="steps between spatial re-sorts (250: 20 equal launches of the 5000-step walk)")
    ap.add_argument("--cpu-sample", type=int, default=200000,
                    help="spins of the CPU-baseline subsample (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--forms", type=int, default=1,
                    help="also measure the batched and served forms (scaled; 0 = skip)")
    ap.add_argument("--forms-steps", type=float, default=1.0,
                    help="step-count scale of the batched/served workloads (1.0 = as stated)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the served stream (the stated stream is 120 s)")
    ap.add_argument("--serve-serial", type=int, default=1,
                    help="also run the served stream's requests one at a time (0 = skip: "
                         "about four times the served makespan)")
    ap.add_argument("--precision-spins", type=int, default=4_000_000)
    ap.add_argument("--precision-steps", type=int, default=10000,
                    help="steps of the precision-form measurement (1e4 = the full workload)")
    ap.add_argument("--precision-fp32-steps", type=int, default=1024,
                    help="steps of the FP32 vector-ALU contraction run (signal comparison)")
    ap.add_argument("--precision-waveforms", type=int, default=1024)
    ap.add_argument("--only", choices=("precision", "batching", "serving"), default=None,
                    help="measure only this form and print its JSON object")
    ap.add_argument("--workload", choices=("standard", "small"), default="standard",
                    help="small: the CPU-runnable 16-axon configuration, GPU and FP64 oracle in full")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only copy a buffer of this many MiB 5 times (HBM counter calibration)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain device copy of known size: the FETCH_SIZE / WRITE_SIZE counters
    of this kernel give the counters' scale (bytes read and written are each
    mib MiB per dispatch)."""
    import torch
    n = int(mib * 2**20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "dispatches": 5}), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only (the walker is "
                                   "single-GPU; spins would shard by range); run with --gpus 1"}))
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    import torch
    if args.workload == "small":
        print(json.dumps(small_workload(args)), flush=True)
        return
    if args.only:
        if args.only == "precision":
            from nozomi_amd.substrate import generate_beaded_axons, load_geometry
            M, L = generate_beaded_axons(args.axons, args.box, seed=args.seed)
            geom = load_geometry(M, L, D=2.0, T2=100.0)
            sub = geom.device_substrate(cell_size=args.cell, margin=args.margin,
                                        device=torch.device("cuda:0"))
            print(json.dumps(precision_form(args, geom, sub)), flush=True)
        else:
            print(json.dumps((batching_form if args.only == "batching" else serving_form)(args)),
                  flush=True)
        return

    from nozomi_amd import runtime as rt
    from nozomi_amd.protocol import PGSEShells, fibonacci_hemisphere
    from nozomi_amd.substrate import generate_beaded_axons, load_geometry

    dev = torch.device("cuda:0")
    M, L = generate_beaded_axons(args.axons, args.box, seed=args.seed)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    prot = PGSEShells(10.0, 40.0, 0.01, [1.0, 2.0, 3.0], fibonacci_hemisphere(64))
    T = min(args.nsteps, prot.nsteps)
    N = args.spins
    sub = geom.device_substrate(cell_size=args.cell, margin=args.margin, device=dev)
    info = sub.info()
    walker = rt.Walker(sub, N, sort_interval=args.sort)
    walker.time_kernels(True)
    start = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=args.seed + 1)
    g = torch.from_numpy(prot.g[:T].copy()).to(dev)
    ga = torch.from_numpy(prot.garea[:T].copy()).to(dev)
    dirs = torch.from_numpy(prot.dirs.astype(np.float32)).to(dev)
    scales = torch.from_numpy(prot.scales.astype(np.float32)).to(dev)
    mom = torch.empty((3, N), dtype=torch.float32, device=dev)
    sums = torch.empty((prot.nsignals, 2), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    walk_ms = []
    launch_ms = []

    def one_pass(record):
        walker.set_positions(start)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        walker.walk(args.seed, T, prot.dt, g=g, garea=ga, t2=False)
        e1.record(stream)
        walker.read(mom=mom)
        rt.signal_sums(mom, dirs, scales, out=sums)
        ms, launches = walker.kernel_time()
        if record:
            walk_ms.append((e0, e1))
            launch_ms.append(ms / launches)

    for _ in range(args.warmup):
        one_pass(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_pass(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    walk_ms = [a.elapsed_time(b) for a, b in walk_ms]
    signals = (sums[:, 0] / N).cpu().numpy()
    gpu_final = None
    if args.cpu_sample > 0:  # the last pass's end state, for the check against the oracle
        n = min(args.cpu_sample, N)
        pos_f64, disp_f64 = walker.read_f64()
        gpu_final = (pos_f64[:, :n], disp_f64[:, :n], mom[:, :n].cpu().numpy().astype(np.float64))

    spin_steps = float(N) * T * args.steps
    value = spin_steps / elapsed
    kern_ms = float(np.mean(walk_ms))

    # roofline of the walk kernel (dominant): FP32 vector work per spin-step
    # (1 + s) * C * 56 + 40 with C = mean candidates per cell, s = mean
    # scatter sub-steps per step (measured below on a subsample); one launch
    # advances N spins by sort_interval steps
    sbar = measure_sbar(geom, sub, prot, dev, args.seed)
    cbar = info["ncand_mean_box"]
    flop_per = (1.0 + sbar) * cbar * 56.0 + 40.0
    launch_avg_ms = float(np.mean(launch_ms))
    steps_per_launch = min(args.sort, T) if args.sort > 0 else T
    launch_units = float(N) * steps_per_launch
    achieved_tf = launch_units * flop_per / (launch_avg_ms * 1e-3) / 1e12
    hbm_gbs = launch_units * 24.0 / (launch_avg_ms * 1e-3) / 1e9
    traffic, traffic_note = measured_traffic(N, steps_per_launch, info)
    # VALU issue: the counted wavefront VALU instructions of a launch (same
    # committed summary, same sources) against the issue limit of one wave64
    # VALU instruction per CU per cycle (no packed FP32)
    valu = measured_traffic(N, steps_per_launch, info, key="valu_insts_per_launch")[0]
    issue_frac = (valu / (launch_avg_ms * 1e-3 * VALU_ISSUE_PER_S) if valu else None)

    out = {
        "metric": "spin-steps/s at 1e6 spins",
        "value": value,
        "unit": "spin-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 state / fp32 scan",
        "data": "synthetic (seeded beaded-axon substrate, uniform spins)",
        "config": {
            "workload": "plain walker FP32 + PGSE phase: 200 axons, 60 um box, 1e6 spins, "
                        "5000 steps, 192 signals",
            "spins": N, "time_steps": T, "dt_ms": prot.dt, "signals": prot.nsignals,
            "spheres": int(info["nspheres"]), "structures": int(info["nstructures"]),
            "cells": list(info["ncell"]), "cell_um": info["cell"][0], "margin_um": info["margin"],
            "mean_candidates": cbar, "scatter_per_step": sbar, "sort_interval": args.sort,
            "precision": "FP64 position, jump and deciding roots in the reference's frame and "
                         "operation order (trajectories bit-identical to the FP64 oracle); FP32 "
                         "candidate scan; FP32 phase moment; FP64 signal sums",
        },
        "walk_ms_per_pass": kern_ms,
        "walk_spin_steps_per_s": float(N) * T / (kern_ms * 1e-3),
        "walk_kernel_launch_ms": launch_avg_ms,
        "signals_per_s": prot.nsignals * args.steps / elapsed,
        "roofline": {
            "bound": "valu", "achieved": achieved_tf, "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved_tf / FP32_VECTOR_PEAK_TFLOPS,
            "traffic": traffic, "traffic_source": traffic_note, "kernel": "walk_kernel",
            "flop_per_spin_step": flop_per,
            "spin_steps_per_launch": launch_units,
            "valu_insts_per_launch": valu, "issue_frac": issue_frac,
        },
        "roofline_hbm": {
            "bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": hbm_gbs / HBM_PEAK_GBS, "bytes_per_spin_step": 24,
        },
        "signal_check": {"b1_mean": float(signals[:64].mean()), "b3_mean": float(signals[128:].mean())},
    }
    if args.cpu_sample > 0:
        out["cpu_baseline"], orc_final = cpu_baseline(geom, prot, start, args)
        if T == min(args.cpu_steps, prot.nsteps):
            out["oracle_check"] = oracle_check(gpu_final, orc_final, prot, geom.Lx, T)
        out["oracle_check_1000_steps"] = trajectory_check(geom, sub, prot, start, args)
    if args.forms:
        out["forms"] = {"precision": precision_form(args, geom, sub)}
        del walker, sub
        torch.cuda.empty_cache()
        out["forms"].update(forms(args))
    print(json.dumps(out), flush=True)


def precision_form(args, geom, sub):
    """W waveform shapes contracted with K-step blocks of displacements (split
    hi/lo BF16 MFMA, FP32 accumulation): K = 64 (the stated configuration) and
    K = 128 over the whole walk, the contraction of block b overlapping the
    walk of block b + 1; each contraction also timed alone on the first
    blocks.  The FP32 vector-ALU contraction runs a shorter walk for the
    signal comparison."""
    import torch
    from nozomi_amd.precision import PrecisionRun, precision_waveforms
    N, T, W, dt = args.precision_spins, args.precision_steps, args.precision_waveforms, 0.01
    g, u, kinds = precision_waveforms(W, T, dt, seed=args.seed)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=args.seed + 3)
    res = {"spins": N, "steps": T, "waveforms": W, "dt_ms": dt,
           "workload": "200-axon substrate, 4e6 spins, 1e4 steps, 1024 PGSE/OGSE waveforms "
                       "(25-150 Hz, |G| <= 300 mT/m, seeded directions)"}

    def timed(K, mode, steps, overlap, nblocks_timed=None):
        pr = PrecisionRun(sub, x0, args.seed, g, u, dt, K=K, mode=mode, overlap=overlap)
        pr.reset()
        pr.block(0)  # warm-up block (kernels, sort)
        pr.reset()
        torch.cuda.synchronize()
        if not overlap:
            pr.gemm_events = []
        t0 = time.perf_counter()
        for b0 in range(0, steps, K):
            pr.block(b0)
        s = pr.signals()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms = None
        if not overlap:
            full = [a.elapsed_time(b) for (a, b), b0 in zip(pr.gemm_events, range(0, steps, K))
                    if b0 + K <= steps]
            ms = float(np.mean(full[:nblocks_timed] if nblocks_timed else full))
        pr.close()
        del pr
        torch.cuda.empty_cache()
        return el, s, ms

    peak_mfma, peak_hbm = 2500.0, HBM_PEAK_GBS
    prog = _progress("precision")
    for K in (64, 128):
        # the contraction alone (one stream) on the first 1280 steps
        nt = min(T, 1280)
        _, _, ms = timed(K, "mfma", nt, overlap=False)
        prog(K, 128)
        el, sig, _ = timed(K, "mfma", T, overlap=True)
        prog(K, 128)
        useful = 6.0 * W * N * K            # FLOP of one block (3 MAC per waveform per step)
        executed = 3.0 * useful             # hi.hi + hi.lo + lo.hi
        # HBM floor of a block: the record read (3K FP32 per spin) and the phase
        # read-modify-write (2 x W FP32 per spin)
        hbm_bytes = N * (12.0 * K + 8.0 * W)
        res[f"K{K}"] = {
            "K": K, "seconds": el, "spin_steps_per_s": N * T / el, "signals_per_s": W / el,
            "contraction_ms_per_block": ms,
            "contraction_useful_tflops": useful / (ms * 1e-3) / 1e12,
            "roofline": {"bound": "mfma", "achieved": executed / (ms * 1e-3) / 1e12,
                         "peak": peak_mfma, "unit": "TFLOP/s",
                         "frac": executed / (ms * 1e-3) / 1e12 / peak_mfma},
            "roofline_hbm": {"bound": "hbm", "bytes_per_block": hbm_bytes,
                             "bytes_per_spin_step": hbm_bytes / (N * K),
                             "achieved": hbm_bytes / (ms * 1e-3) / 1e9, "peak": peak_hbm,
                             "unit": "GB/s", "frac": hbm_bytes / (ms * 1e-3) / 1e9 / peak_hbm,
                             "floor_ms_per_block": hbm_bytes / (peak_hbm * 1e9) * 1e3},
        }
        res[f"K{K}"]["_signals"] = sig
    # FP32 vector-ALU contraction against the MFMA one on the same shorter walk
    nt = min(T, args.precision_fp32_steps)
    _, s_f32, ms32 = timed(128, "fp32", nt, overlap=False)
    _, s_mfma, _ = timed(128, "mfma", nt, overlap=False)
    res["fp32_contraction"] = {"steps": nt, "K": 128, "contraction_ms_per_block": ms32,
                               "useful_tflops": 6.0 * W * N * 128 / (ms32 * 1e-3) / 1e12}
    res["max_signal_diff_mfma_vs_fp32"] = float(np.abs(s_mfma - s_f32).max())
    res["max_signal_diff_K64_vs_K128"] = float(np.abs(res["K64"].pop("_signals") -
                                                      res["K128"].pop("_signals")).max())
    res["note_K64_vs_K128"] = ("K = 128 accumulates its two 64-step parts separately and adds "
                               "them to the phases in order: the same FP32 operations as two "
                               "K = 64 blocks, so the signals agree bit for bit")
    return res


def small_workload(args):
    """The CPU-runnable configuration in full on the GPU and with the FP64
    oracle: 16 beaded axons in a 30 um periodic cube, 1e4 spins uniform in the
    box, D = 2.0 um^2/ms, dt = 20 us, PGSE delta = 10 ms, Delta = 30 ms
    (2000 steps), b = 1 ms/um^2, 6 directions."""
    import torch
    from nozomi_amd import runtime as rt
    from nozomi_amd.protocol import PGSEShells, icosahedron_6
    from nozomi_amd.substrate import generate_beaded_axons, load_geometry
    from oracle.oracle import OracleSubstrate, build
    dev = torch.device("cuda:0")
    M, L = generate_beaded_axons(16, 30.0, seed=args.seed)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    prot = PGSEShells(10.0, 30.0, 0.02, [1.0], icosahedron_6())
    N, T = 10000, prot.nsteps
    sub = geom.device_substrate(device=dev)
    start = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=args.seed + 1)
    w = rt.Walker(sub, N, sort_interval=args.sort)
    g = torch.from_numpy(prot.g.copy()).to(dev)
    ga = torch.from_numpy(prot.garea.copy()).to(dev)
    dirs = torch.from_numpy(prot.dirs.astype(np.float32)).to(dev)
    scales = torch.from_numpy(prot.scales.astype(np.float32)).to(dev)
    mom = torch.empty((3, N), dtype=torch.float32, device=dev)
    times = []
    for rep in range(4):  # the first pass warms up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        w.set_positions(start)
        w.walk(args.seed, T, prot.dt, g=g, garea=ga, t2=False)
        w.read(mom=mom)
        sums = rt.signal_sums(mom, dirs, scales).cpu().numpy()
        times.append(time.perf_counter() - t0)
    sg = (sums[:, 0] + 1j * sums[:, 1]) / N
    pos_g, disp_g = w.read_f64()
    w.close()
    build()
    orc = OracleSubstrate(geom.spheres, geom.nstructures, geom.structure_D(), geom.structure_T2(),
                          (L, L, L))
    p = start.cpu().numpy().astype(np.float64)
    pos, pos0, sigw, mo = p.copy(), p.copy(), np.ones(N), np.zeros((3, N))
    threads = min(len(os.sched_getaffinity(0)), 16)
    t0 = time.perf_counter()
    orc.walk(pos, pos0, sigw, seed=args.seed, step0=0, nsteps=T, dt=prot.dt,
             g=prot.g.astype(np.float64), garea=prot.garea.astype(np.float64), mom=mo,
             nthreads=threads)
    t_cpu = time.perf_counter() - t0
    so = np.exp(1j * prot.scales[None, :] * (mo.T @ prot.dirs.astype(np.float64))).mean(0)
    d = pos_g - pos
    d -= L * np.round(d / L)
    err = np.maximum(np.abs(d).max(0), np.abs(disp_g - (pos - pos0)).max(0))
    t_gpu = float(np.median(times[1:]))
    return {
        "workload": "16 axons, 30 um box, 1e4 spins, 2000 steps of 20 us, PGSE 10/30 ms, "
                    "b = 1 ms/um^2, 6 directions",
        "gpu_seconds": t_gpu, "gpu_spin_steps_per_s": N * T / t_gpu,
        "cpu_seconds": t_cpu, "cpu_spin_steps_per_s": N * T / t_cpu, "cpu_threads": threads,
        "cpu_kind": "port (FP64 oracle, OpenMP)",
        "signals_gpu": [float(x) for x in sg.real], "signals_oracle": [float(x) for x in so.real],
        "signal_max_abs_diff": float(np.abs(sg - so).max()),
        "signal_sampling_se": float(np.sqrt(np.maximum(1 - np.abs(so) ** 2, 0) / N).mean()),
        "free_diffusion_exp_minus_bD": float(np.exp(-prot.b_per_signal[0] * 2.0)),
        "trajectory_fraction_within_1e-5_um": float(np.mean(err <= 1e-5)),
        "trajectory_error_median_um": float(np.median(err)),
    }


def _progress(what):
    """Progress lines on stderr for the long form measurements."""
    def report(done, total):
        print(f"[{what}] {done}/{total} done", file=sys.stderr, flush=True)
    return report


def forms(args):
    """Batched and served forms against back-to-back plain walker launches of
    the same simulations (each result checked against its run alone)."""
    return {"batching": batching_form(args), "serving": serving_form(args)}


WALK_LANES = 256 * 4 * 4 * 64  # resident lanes of the walk kernel (4 waves per SIMD)


def full_occupancy_rate(reqs, cache, seed, nsample=8, n=2_000_000):
    """Spin-steps/s of a walk that fills the GPU (n spins: about eight rounds
    of the resident lanes) on a sample of the mix's substrates, over two of
    their tuned sort intervals: what any schedule of the mix can reach at
    most per spin-step.  Returns (max, mean) over the sample."""
    import torch
    from nozomi_amd import runtime as rt
    rng = np.random.default_rng(seed)
    keys = list(dict.fromkeys(r.key for r in reqs))
    rates = []
    for i in rng.choice(len(keys), size=min(nsample, len(keys)), replace=False):
        r = next(q for q in reqs if q.key == keys[i])
        sub, geom = cache.acquire(r.key, r.make_geometry)
        k = cache.params(r.key)["k"] or 256
        steps = 2 * k
        w = rt.Walker(sub, n, sort_interval=k)
        w.seed(targets=np.arange(geom.nstructures + 1), seed=seed + int(i))
        w.walk(seed, k, r.dt, t2=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        w.walk(seed, steps, r.dt, t2=False)
        torch.cuda.synchronize()
        rates.append(n * steps / (time.perf_counter() - t0))
        w.close()
        cache.release(r.key)
    return float(max(rates)), float(np.mean(rates))


def batching_form(args):
    import torch
    from nozomi_amd.serve import SubstrateCache, run_batched, run_serial
    from nozomi_amd.workloads import batching_mix, pregenerate, total_spin_steps
    dev = torch.device("cuda:0")
    # batching: 256 independent simulations of uneven size in one pool; their
    # substrates are generated, tuned and uploaded before timing (resident)
    reqs = batching_mix(nsims=256, seed=args.seed, scale_steps=args.forms_steps)
    pregenerate(reqs)
    cache = SubstrateCache(dev, 1 << 40, tune=True)
    cache.prefetch(reqs)
    from nozomi_amd.autotune import family_key
    fam = {}
    for v in cache.items.values():
        fam.setdefault(str(family_key(v[1])), {"cell": v[4]["cell"], "margin": v[4]["margin"],
                                               "k": v[4]["k"], "substrates": 0})
        fam[str(family_key(v[1]))]["substrates"] += 1
    cap = sum(-(-r.n // 256) * 256 for r in reqs)
    run_serial(reqs[:4], cache=cache)  # warms the kernels
    got, t_b = run_batched(reqs, capacity=cap, k=96, cache=cache, progress=_progress("batched"))
    alone, t_s = run_serial(reqs, cache=cache, progress=_progress("serial"))
    diff = max(float(np.abs(got[r].signals - alone[r].signals).max()) for r in got)
    ss = total_spin_steps(reqs)
    # what any schedule could gain: each simulation at the full-occupancy rate
    rmax, rmean = full_occupancy_rate(reqs, cache, args.seed)
    dur = np.array([alone[r.rid].latency for r in reqs])
    work = np.array([float(r.spin_steps) for r in reqs])
    lanes = WALK_LANES
    under = np.array([r.n < lanes for r in reqs])
    return {
        "sims": len(reqs), "spin_steps": ss, "step_scale": args.forms_steps,
        "batched_sims_per_s": len(reqs) / t_b, "serial_sims_per_s": len(reqs) / t_s,
        "speedup_vs_serial": t_s / t_b, "batched_spin_steps_per_s": ss / t_b,
        "serial_spin_steps_per_s": ss / t_s, "max_signal_diff_vs_alone": diff,
        "autotuned": fam,
        "amdahl": {
            "full_occupancy_spin_steps_per_s_max": rmax,
            "full_occupancy_spin_steps_per_s_mean": rmean,
            "speedup_bound": float(dur.sum() / (work.sum() / rmax)),
            "speedup_bound_at_mean_rate": float(dur.sum() / (work.sum() / rmean)),
            "serial_seconds": float(dur.sum()),
            "serial_share_of_underfilled_sims": float(dur[under].sum() / dur.sum()),
            "underfilled_sims": int(under.sum()), "underfilled_below_spins": lanes,
            "note": "bound = serial time / (spin-steps / best full-occupancy rate): no "
                    "schedule walks a spin-step faster than a GPU-filling plain walk on the "
                    "fastest sampled substrate; at_mean_rate uses the sample's mean rate "
                    "(every substrate walked at the full-occupancy rate typical of the mix)",
        },
    }


def serving_form(args):
    from nozomi_amd.serve import (Server, fifo_latencies, latency_stats, run_serial, serve_trace)
    from nozomi_amd.workloads import pregenerate, serving_mix, total_spin_steps
    from nozomi_amd.serve import SubstrateCache
    import torch
    dev = torch.device("cuda:0")
    # serving: Poisson arrivals at 50 req/s, continuous batching, HBM admission;
    # device substrates are built when a request first needs one.  The
    # substrate families are tuned before the trace (deployment time); the
    # served and the serial leg each start from an empty substrate cache and
    # build what their requests need.
    reqs, arr = serving_mix(rate=50.0, duration=args.serve_seconds, seed=args.seed + 1,
                            scale_steps=args.forms_steps)
    pregenerate(reqs)
    t0 = time.perf_counter()
    families = SubstrateCache(dev, 0, tune=True).pretune(reqs)
    t_tune = time.perf_counter() - t0
    run_serial(reqs[:2], cache=SubstrateCache(dev, 1 << 40, tune=True))  # warms the kernels
    cap = 16 * 1024 * 1024
    srv = Server(cap, hbm_budget=64 << 30, max_sims=256, k=64, tune=True)
    got, makespan = serve_trace(srv, reqs, arr, progress=_progress("served"))
    built, reused = srv.subs.built, srv.subs.reused
    srv.close()
    ss = total_spin_steps(reqs)
    if args.serve_serial:
        alone, t_s = run_serial(reqs, cache=SubstrateCache(dev, 64 << 30, tune=True),
                                progress=_progress("serial"))
        diff = max(float(np.abs(got[r].signals - alone[r].signals).max()) for r in got)
        lat_serial = fifo_latencies(arr, [alone[r.rid].latency for r in reqs])
        serial = {"serial_sims_per_s": len(reqs) / t_s,
                  "serial_fifo_latency": {"p50_s": float(np.percentile(lat_serial, 50)),
                                          "p99_s": float(np.percentile(lat_serial, 99))},
                  "max_signal_diff_vs_alone": diff}
    else:
        serial = {"serial_sims_per_s": None, "serial_fifo_latency": None,
                  "max_signal_diff_vs_alone": None}
    return {
        "requests": len(reqs), "rate_per_s": 50.0, "arrival_window_s": args.serve_seconds,
        "step_scale": args.forms_steps, "offered_spin_steps_per_s": ss / args.serve_seconds,
        "makespan_s": makespan, "served_sims_per_s": len(reqs) / makespan,
        "served_spin_steps_per_s": ss / makespan, **serial,
        "latency": latency_stats(got),
        "substrates_built": built, "substrates_reused": reused,
        "families_tuned": len(families), "tuning_s_before_trace": t_tune,
    }


WALK_SOURCES = ("nz_walk.hip", "nz_walk_core.h", "nz_internal.h", "nz_rng.h", "nz_grid_build.h")


def walk_source_hash():
    """sha256 (first 16 hex digits) of the walk kernel's sources."""
    import hashlib
    h = hashlib.sha256()
    for name in WALK_SOURCES:
        with open(os.path.join(ROOT, "nozomi_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(n, steps_per_launch, info, key="bytes_per_launch"):
    """HBM bytes per walk-kernel launch from the committed counter summary
    (profiles/walk_kernel_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by the copy-kernel calibration), or another of its figures (key),
    when it was measured on this workload shape and on the walk kernel built
    from the current sources; else None.  Returns (value or None, why)."""
    path = os.path.join(ROOT, "profiles", "walk_kernel_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, "no counter summary"
    if (t.get("spins") != n or t.get("steps_per_launch") != steps_per_launch
            or t.get("cells") != list(info["ncell"])):
        return None, "counter summary measured on another workload shape"
    if t.get("sources_sha16") != walk_source_hash():
        return None, "counter summary measured on other walk-kernel sources"
    return t.get(key), t.get("source")


def measure_sbar(geom, sub, prot, dev, seed, n=65536, steps=200):
    """Mean scatter sub-steps per spin-step on a subsample (GPU counter)."""
    from nozomi_amd import runtime as rt
    w = rt.Walker(sub, n, counters=True)
    w.seed(targets=np.arange(geom.nstructures + 1), seed=seed + 7)
    w.walk(seed, steps, prot.dt, t2=False)
    nsub, _ = w.counters()
    return float(nsub.sum()) / (n * steps)


def cpu_baseline(geom, prot, start, args):
    """FP64 oracle (OpenMP over spins) on the first cpu_sample spins for
    cpu_steps steps of the same workload."""
    from oracle.oracle import OracleSubstrate, build
    build()
    n = min(args.cpu_sample, start.shape[1])
    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 16)
    orc = OracleSubstrate(geom.spheres, geom.nstructures, geom.structure_D(), geom.structure_T2(),
                          (geom.Lx, geom.Ly, geom.Lz))
    p = start[:, :n].cpu().numpy().astype(np.float64)
    pos, pos0, sig, mom = p.copy(), p.copy(), np.ones(n), np.zeros((3, n))
    T = min(args.cpu_steps, prot.nsteps)
    t0 = time.perf_counter()
    orc.walk(pos, pos0, sig, seed=args.seed, step0=0, nsteps=T, dt=prot.dt,
             g=prot.g[:T].astype(np.float64), garea=prot.garea[:T].astype(np.float64), mom=mom,
             nthreads=threads)
    el = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return ({"value": n * T / el, "unit": "spin-steps/s", "cores": threads, "kind": "port",
             "sample": f"{n} spins x {T} steps of the same workload (FP64 oracle, OpenMP)",
             "cpu_model": model, "host_cores_available": cores}, (pos, pos0, mom))


def oracle_check(gpu, orc, prot, L, T):
    """Errors of the GPU walk against the FP64 oracle on the oracle's sample
    (same spins, start positions and Philox stream): per-spin trajectory
    error (wrapped position and unwrapped displacement, um) and the signals
    of the protocol formed from each side's phase moments."""
    pos_g, disp_g, mom_g = gpu
    pos_o, pos0_o, mom_o = orc
    d = pos_g - pos_o
    d -= L * np.round(d / L)
    err = np.maximum(np.abs(d).max(0), np.abs(disp_g - (pos_o - pos0_o)).max(0))
    dirs = prot.dirs.astype(np.float64)
    scales = prot.scales.astype(np.float64)

    def sig(m):
        return np.exp(1j * scales[None, :] * (m.T @ dirs)).mean(0)

    sg, so = sig(mom_g), sig(mom_o)
    ds = np.abs(sg - so)
    # standard error of one side's sample mean of exp(i phase), for scale
    se = np.sqrt(np.maximum(1.0 - np.abs(so) ** 2, 0.0) / err.size)
    return {
        "spins": int(err.size), "steps": int(T),
        "trajectory_error_um": {"median": float(np.median(err)),
                                "p90": float(np.percentile(err, 90)),
                                "p99": float(np.percentile(err, 99)), "max": float(err.max())},
        "fraction_within_1e-5_um": float(np.mean(err <= 1e-5)),
        "fraction_bit_identical": float(np.mean(err == 0.0)),
        "signal_max_abs_diff": float(ds.max()), "signal_mean_abs_diff": float(ds.mean()),
        "signal_bar_1e-3_met": bool(ds.max() < 1e-3),
        "signal_sampling_se_mean": float(se.mean()),
        "note": "GPU (FP64 state, FP32 scan) against the FP64 oracle on the same spins, start "
                "positions and Philox stream, all 5000 steps; the signals differ only by the "
                "GPU's FP32 phase-moment accumulation",
    }


def trajectory_check(geom, sub, prot, start, args, n=20000, T=1000):
    """Per-spin trajectories of n spins over T steps, GPU walker against the
    FP64 oracle (same start positions and Philox stream), split by the number
    of scatter sub-steps each spin took in the oracle's walk."""
    import torch
    from nozomi_amd import runtime as rt
    from oracle.oracle import OracleSubstrate, build
    build()
    n = min(n, start.shape[1])
    T = min(T, prot.nsteps)
    x0 = start[:, :n].contiguous()
    w = rt.Walker(sub, n, sort_interval=args.sort)
    w.set_positions(x0)
    w.walk(args.seed, T, prot.dt, t2=False)
    pos_g, disp_g = w.read_f64()
    w.close()
    orc = OracleSubstrate(geom.spheres, geom.nstructures, geom.structure_D(), geom.structure_T2(),
                          (geom.Lx, geom.Ly, geom.Lz))
    p = x0.cpu().numpy().astype(np.float64)
    pos, pos0, sigw = p.copy(), p.copy(), np.ones(n)
    nsub = np.zeros(n, np.int64)
    orc.walk(pos, pos0, sigw, seed=args.seed, step0=0, nsteps=T, dt=prot.dt, nsub=nsub,
             nthreads=min(len(os.sched_getaffinity(0)), 16))
    L = geom.Lx
    d = pos_g - pos
    d -= L * np.round(d / L)
    err = np.maximum(np.abs(d).max(0), np.abs(disp_g - (pos - pos0)).max(0))
    few = nsub <= 60
    return {
        "spins": int(n), "steps": int(T),
        "fraction_within_1e-5_um": float(np.mean(err <= 1e-5)),
        "fraction_bit_identical": float(np.mean(err == 0.0)),
        "spins_with_at_most_60_scatter_substeps": int(few.sum()),
        "fraction_within_1e-5_um_at_most_60": float(np.mean(err[few] <= 1e-5)),
        "median_error_um_at_most_60": float(np.median(err[few])),
        "median_scatter_substeps": float(np.median(nsub)),
    }


if __name__ == "__main__":
    main()
