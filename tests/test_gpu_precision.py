"""The precision form: K-step blocks of displacements contracted with many
waveforms as split hi/lo BF16 MFMA products (FP32 accumulation), against an
FP64 contraction of the same blocks, the FP32 vector-ALU contraction, and the
FP64 oracle's own walk of the same spins."""
import numpy as np
import pytest
import torch

from nozomi_amd._lib import check, lib
from nozomi_amd.precision import PrecisionRun, precision_waveforms
from nozomi_amd.runtime import _ptr, _stream_ptr
from nozomi_amd.substrate import generate_beaded_axons, load_geometry

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.mark.parametrize("W", [200, 24])
@pytest.mark.parametrize("K", [32, 64, 96, 128])
def test_gemm_modes_against_fp64(K, W):
    """200 waveforms: seven 32-waveform tiles, the last part-filled; 24: one
    tile (its epilogue only after the tile loop).  1000 slots: the last
    128-slot workgroup part-filled, every seventh slot padding."""
    rng = np.random.default_rng(K + W)
    N = 1000
    rec = rng.normal(scale=5.0, size=(N, 3 * K)).astype(np.float32)  # slot-major
    bt = rng.normal(scale=1e-3, size=(W, 3 * K)).astype(np.float32)
    slot_spin = rng.permutation(N).astype(np.int32)
    slot_spin[::7] = -1                      # padding slots contribute nothing
    ph0 = rng.normal(size=(N, W)).astype(np.float32)
    ref = ph0.astype(np.float64).copy()
    valid = slot_spin >= 0
    ref[slot_spin[valid]] += rec[valid].astype(np.float64) @ bt.T.astype(np.float64)
    # BF16 hi/lo of bt the way the library splits (round to nearest even)
    u = bt.view(np.uint32).astype(np.uint64)
    hi = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    lo_f = bt - (hi.astype(np.uint32) << 16).view(np.float32)
    ul = lo_f.view(np.uint32).astype(np.uint64)
    lo = ((ul + 0x7FFF + ((ul >> 16) & 1)) >> 16).astype(np.uint16)
    t = {k: torch.from_numpy(v).to(DEV) for k, v in
         dict(rec=rec, ss=slot_spin, hi=hi.view(np.int16), lo=lo.view(np.int16), f32=bt).items()}
    scale = np.abs(rec).astype(np.float64) @ np.abs(bt).T.astype(np.float64)
    for mode, tol in ((0, 6e-5), (1, 1e-5)):
        ph = torch.from_numpy(ph0.copy()).to(DEV)
        check(lib().nz_phase_gemm(_ptr(t["rec"]), 3 * K, N, _ptr(t["ss"]), K, _ptr(t["hi"]), _ptr(t["lo"]),
                                  _ptr(t["f32"]), W, mode, _ptr(ph), _stream_ptr(DEV)))
        got = ph.cpu().numpy().astype(np.float64)
        err = np.abs(got - ref)
        rel = err[slot_spin[valid]] / (scale[valid] + 1e-30)
        print(f"mode {mode}: max error relative to sum |terms| = {rel.max():.2e}")
        assert rel.max() < tol
        untouched = np.setdiff1d(np.arange(N), slot_spin[valid])
        assert np.array_equal(got[untouched], ph0[untouched].astype(np.float64))


def test_waveform_count_padded_to_four():
    """W = 50 (not a multiple of 4): the run pads two zero waveforms for the
    MFMA contraction and reports 50 signals, equal to the FP32 contraction's
    within the split-BF16 bound."""
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N, T, dt, W, K = 8192, 96, 0.02, 50, 32
    g, u, _ = precision_waveforms(W, T, dt, seed=5)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=3)
    out = {}
    for mode in ("mfma", "fp32"):
        pr = PrecisionRun(sub, x0, 17, g, u, dt, K=K, mode=mode)
        pr.run()
        out[mode] = pr.signals()
        assert pr.W == 52 and out[mode].shape == (W,)
        pr.close()
    d = np.abs(out["mfma"] - out["fp32"]).max()
    print(f"W = 50: max |signal MFMA - FP32| = {d:.1e}")
    assert d < 1e-5


def test_precision_walk_against_oracle(oracle_lib):
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N, T, dt, W, K, seed = 4096, 256, 0.02, 48, 64, 4321
    g, u, kinds = precision_waveforms(W, T, dt, seed=7)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=11)
    blocks = []

    def hook(t0, k, rec, slot_spin):
        blocks.append((t0, k, rec[:, :3 * k].cpu().numpy().copy(), slot_spin.cpu().numpy().copy()))

    runs = {}
    for mode in ("mfma", "fp32"):
        pr = PrecisionRun(sub, x0, seed, g, u, dt, K=K, mode=mode)
        pr.run(record_hook=hook if mode == "mfma" else None)
        pr.finish()  # the contraction stream is done before the phases are read
        runs[mode] = (pr.phase.cpu().numpy().astype(np.float64), pr.signals())
        pr.close()
    # FP64 contraction of the recorded FP32 displacements: contraction error alone
    x0h = x0.cpu().numpy().astype(np.float64)
    g64, u64 = g.astype(np.float64), u.astype(np.float64)
    ph_rec = (x0h.T @ u64) * g64.sum(1)[None, :]
    for t0, k, rec, ss in blocks:
        d = rec.reshape(N, k, 3).transpose(1, 2, 0)                # [t, a, slot]
        contrib = np.einsum("tas,wt,aw->sw", d, g64[:, t0:t0 + k], u64)
        ph_rec[ss] += contrib
    for mode in ("mfma", "fp32"):
        e = np.abs(runs[mode][0] - ph_rec).max()
        print(f"{mode}: max |phase - FP64 contraction of the record| = {e:.2e} rad")
        assert e < (2e-4 if mode == "mfma" else 2e-5)
    # the FP64 oracle walking the same spins: phases from its own trajectories
    orc = oracle_lib.OracleSubstrate(geom.spheres, geom.nstructures, geom.structure_D(),
                                     geom.structure_T2(), (L, L, L))
    pos, pos0 = x0h.copy(), x0h.copy()
    ph_o = ph_rec * 0 + (x0h.T @ u64) * g64.sum(1)[None, :]
    for t in range(T):
        orc.walk(pos, pos0, np.ones(N), seed=seed, step0=t, nsteps=1, dt=dt)
        ph_o += ((pos - pos0).T @ (u64 * g64[:, t][None, :]))
    s_o = np.exp(1j * ph_o).mean(0)
    for mode in ("mfma", "fp32"):
        es = np.abs(runs[mode][1] - s_o).max()
        print(f"{mode}: max |signal - FP64 oracle| = {es:.2e}")
        assert es < 1e-3
    assert np.abs(runs["mfma"][1] - runs["fp32"][1]).max() < 1e-4


def test_overlapped_blocks_equal_serial():
    """Contracting block b on a second stream while block b + 1 walks gives
    the phases of the one-stream run, bit for bit (same kernels, same order
    of the phase updates), also when a block ran before the reset."""
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N, T, dt, W, K, seed = 20000, 640, 0.02, 96, 64, 99
    g, u, _ = precision_waveforms(W, T, dt, seed=3)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=4)
    out = {}
    for overlap in (False, True):
        pr = PrecisionRun(sub, x0, seed, g, u, dt, K=K, mode="mfma", overlap=overlap)
        pr.reset()
        pr.block(0)  # its contraction may still run when run() resets
        pr.run()
        pr.finish()
        out[overlap] = (pr.phase.cpu().numpy().copy(), pr.signals())
        pr.close()
    assert np.array_equal(out[False][0], out[True][0])
    assert np.array_equal(out[False][1], out[True][1])


def test_precision_form_full_size_two_blocks():
    """The precision configuration's size: 4e6 spins on the 200-axon
    substrate, 1024 waveforms, two K = 64 blocks.  The phases of a 4096-spin
    subsample equal an FP64 contraction of the same recorded displacements to
    the split-BF16 bound; every phase is finite."""
    M, L = generate_beaded_axons(200, 60.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N, T, dt, W, K, seed = 4_000_000, 128, 0.01, 1024, 64, 77
    g, u, _ = precision_waveforms(W, T, dt, seed=20261203)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=5)
    pick = torch.arange(0, N, N // 4096, device=DEV)[:4096]  # spin ids of the subsample
    blocks = []

    def hook(t0, k, rec, slot_spin):
        slot_of = torch.empty(N, dtype=torch.int64, device=DEV)
        slot_of[slot_spin.long()] = torch.arange(N, device=DEV)
        cols = slot_of[pick]
        blocks.append((t0, k, rec.index_select(0, cols)[:, :3 * k].cpu().numpy().astype(np.float64)))

    print(f"precision form: {N} spins x {T} steps, {W} waveforms, K = {K}", flush=True)
    pr = PrecisionRun(sub, x0, seed, g, u, dt, K=K, mode="mfma")
    pr.run(record_hook=hook)
    pr.finish()
    ph = pr.phase.index_select(0, pick).cpu().numpy().astype(np.float64)
    finite = bool(torch.isfinite(pr.phase).all().item())
    pr.close()
    assert finite and len(blocks) == 2
    g64, u64 = g.astype(np.float64), u.astype(np.float64)
    x0h = x0.index_select(1, pick).cpu().numpy().astype(np.float64)
    ref = (x0h.T @ u64) * g64.sum(1)[None, :]
    scale = (np.abs(x0h).T @ np.abs(u64)) * np.abs(g64).sum(1)[None, :]
    for t0, k, d in blocks:
        d = d.reshape(-1, k, 3).transpose(1, 2, 0)  # [t, a, spin of the subsample]
        ref += np.einsum("tas,wt,aw->sw", d, g64[:, t0:t0 + k], u64)
        scale += np.einsum("tas,wt,aw->sw", np.abs(d), np.abs(g64[:, t0:t0 + k]), np.abs(u64))
    err = np.abs(ph - ref)
    print(f"  max |phase - FP64| = {err.max():.2e} rad; relative to sum |terms| "
          f"{(err / (scale + 1e-30)).max():.2e}", flush=True)
    assert (err / (scale + 1e-30)).max() < 6e-5
    assert err.max() < 2e-3


def test_run_ahead_record_equals_stopping_at_block_ends():
    """Spins whose block is done walk up to `ahead` steps of the next block
    while the slower spins of their wavefront finish (nz_walk_record_ring),
    leaving those steps in the ring for the next block.  Every block's record
    (in spin order) and the phases equal those of the run whose spins all stop
    at each block end, bit for bit: with and without the contraction on its own
    stream, re-sorting every block or every fourth, and with a short last
    block."""
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N, T, dt, W, K, seed = 20000, 520, 0.02, 64, 64, 31  # eight blocks of 64 steps and one of 8
    g, u, _ = precision_waveforms(W, T, dt, seed=9)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=6)
    out = {}
    for ahead, overlap, interval in ((0, False, 64), (64, False, 256), (64, True, 256),
                                     (32, True, 128), (64, True, 64)):
        recs = []

        def hook(t0, k, rec, slot_spin):
            r = torch.empty((N, 3 * k), dtype=torch.float32, device=DEV)
            r[slot_spin.long()] = rec
            recs.append(r.cpu().numpy())

        pr = PrecisionRun(sub, x0, seed, g, u, dt, K=K, mode="mfma", overlap=overlap, ahead=ahead,
                          sort_interval=interval)
        pr.run(record_hook=hook)
        pr.finish()
        out[(ahead, overlap, interval)] = (recs, pr.phase.cpu().numpy().copy(), pr.signals())
        pr.close()
    ref = out[(0, False, 64)]
    assert len(ref[0]) == 9
    for key, (recs, ph, sig) in out.items():
        print(f"ahead {key[0]}, overlap {key[1]}, re-sort every {key[2]} steps: "
              "records and phases equal", flush=True)
        assert len(recs) == len(ref[0])
        for a, b in zip(recs, ref[0]):
            assert np.array_equal(a, b), key
        assert np.array_equal(ph, ref[1]), key
        assert np.array_equal(sig, ref[2]), key


@pytest.mark.parametrize("K", [32, 96, 128])
def test_run_ahead_other_block_lengths(K):
    """The run-ahead ring at the other MFMA block lengths (32-k chunks for
    K = 32 and 96, two parts for K = 128): phases and signals equal stopping
    every spin at each block end, bit for bit, with a short last block."""
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N, dt, W, seed = 12000, 0.02, 40, 8
    T = 4 * K + 12
    g, u, _ = precision_waveforms(W, T, dt, seed=K)
    x0 = sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=K + 1)
    out = []
    for ahead in (0, K // 2):
        pr = PrecisionRun(sub, x0, seed, g, u, dt, K=K, mode="mfma", ahead=ahead)
        pr.run()
        pr.finish()
        out.append((pr.phase.cpu().numpy().copy(), pr.signals()))
        pr.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_run_ahead_record_argument_rules():
    """nz_walk_record_ring refuses, before any launch: a block that wraps the
    ring, block + ahead longer than the ring, rows shorter than the ring, a
    record call after a run-ahead block that does not continue its ring where
    it ended, and a block shorter than the steps run ahead."""
    from nozomi_amd import runtime as rt
    M, L = generate_beaded_axons(6, 20.0, seed=20261203)
    geom = load_geometry(M, L, D=2.0, T2=100.0)
    sub = geom.device_substrate(device=DEV)
    N = 1024
    w = rt.Walker(sub, N, sort_interval=1024)
    w.set_positions(sub.seed(N, targets=np.arange(geom.nstructures + 1), seed=1))
    rec = torch.zeros((N, 3 * 128), dtype=torch.float32, device=DEV)
    other = torch.zeros_like(rec)
    ss = torch.empty(N, dtype=torch.int32, device=DEV)
    Lb, st = lib(), _stream_ptr(DEV)

    def call(nsteps, ahead, pos0, r=rec, ld=3 * 128, ring=128):
        return Lb.nz_walk_record_ring(w.handle, 5, nsteps, ahead, 0.02, _ptr(r), ld, ring, pos0,
                                      _ptr(ss), st)

    def err():
        return Lb.nz_last_error_string()

    assert call(64, 0, 96) < 0 and b"wrapping" in err()
    assert call(64, 65, 0) < 0 and b"exceed the ring" in err()
    assert call(64, 0, 0, ld=3 * 127) < 0 and b"ld" in err()
    assert call(64, 32, 0) == 0                   # runs up to 32 steps ahead
    assert call(64, 0, 0) < 0 and b"entry 64" in err()
    assert call(64, 0, 64, r=other) < 0 and b"same ring" in err()
    assert call(64, 0, 64, ld=3 * 130) < 0 and b"same ring" in err()
    assert Lb.nz_walk_record(w.handle, 5, 64, 0.02, _ptr(rec), 3 * 128, _ptr(ss), st) < 0
    assert call(16, 0, 64) < 0 and b"further ahead" in err()
    assert call(64, 32, 64) == 0                  # continues at entry 64, ahead into entry 0..
    assert call(32, 0, 0) == 0                    # takes those steps over and stops at its end
    assert Lb.nz_walk_record(w.handle, 5, 64, 0.02, _ptr(rec), 3 * 128, _ptr(ss), st) == 0
    torch.cuda.synchronize()
    w.close()
