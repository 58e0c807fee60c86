"""The precision form of the DWI walker: many waveform shapes at once.

W gradient waveforms (PGSE and apodized-cosine OGSE, |G| <= 300 mT/m, seeded
random directions) sampled on the walk's time grid are contracted with the
spins' displacements in blocks of K steps: [N x 3K] x [3K x W] per block on
the MFMA units as split hi/lo BF16 with FP32 accumulation (nz_phase_gemm mode
0), or on the vector ALUs in FP32 (mode 1).  The phases [N, W] stay resident
in HBM; the signals sum exp(i phase) over spins in FP64 at the end.

Phase per step: phi += g_w(t) u_w . x(t) with g_w(t) = gamma G_w(t) dt,
gamma = 2.67513e-4 rad / (ms um mT/m) (267.513 rad/ms/mT, waveforms.py:4).
Positions enter as start position + unwrapped displacement, so no wrap
correction is involved (the plain walker keeps the reference's correction).
"""
import ctypes

import numpy as np
import torch

from . import runtime as rt
from ._lib import check, lib
from .runtime import _ptr, _stream_ptr
from .simulation_engine import waveforms as wf

GAMMA = 2.67513e-4  # rad / (ms um (mT/m))


def precision_waveforms(W, T, dt, seed=20261203, gmax=300.0):
    """(g [W, T] float32 per-step coefficients, u [3, W] float32 unit
    directions, kinds [W] str).  Half PGSE (delta, Delta drawn to fit the
    echo time T dt), half apodized cosine OGSE at 25-150 Hz."""
    rng = np.random.default_rng(seed)
    te = T * dt
    g = np.zeros((W, T), np.float64)
    kinds = []
    t = dt * np.arange(1, T + 1)
    for w in range(W):
        amp = rng.uniform(0.2, 1.0) * gmax
        if w % 2 == 0:
            delta = rng.uniform(0.1, 0.4) * te
            Delta = rng.uniform(delta + dt, te - delta)
            on1 = (t > te / 2 - Delta / 2 - delta / 2) & (t < te / 2 - Delta / 2 + delta / 2)
            on2 = (t > te / 2 + Delta / 2 - delta / 2) & (t < te / 2 + Delta / 2 + delta / 2)
            wave = on1.astype(float) - on2.astype(float)
            kinds.append("PGSE")
        else:
            f = rng.uniform(25.0, 150.0) / 1000.0  # cycles per ms
            n_cyc = max(1, int(0.45 * te * f - 0.25))
            dur = min((0.25 + n_cyc) / f, 0.48 * te)
            o = wf.ApodizedCosineOGSEWaveform(n_cyc, dur, te=te, time_step=dt)
            wave = np.resize(o.wave, T)
            kinds.append("OGSE")
        g[w] = GAMMA * amp * wave * dt
    u = rng.normal(size=(3, W))
    u /= np.linalg.norm(u, axis=0, keepdims=True)
    return g.astype(np.float32), u.astype(np.float32), kinds


class PrecisionRun:
    """Walk N spins (resident walker) for T steps and contract every K-step
    block with the W waveforms.  mode: "mfma" (split BF16) or "fp32".

    overlap: the walk of block b+1 (vector ALUs) runs on the walker's stream
    while the contraction of block b (matrix cores, or the FP32 kernel) runs on
    a second stream; the slot order and the waveform block are double-buffered,
    the displacement record is a ring of three blocks, and events order each
    buffer's reuse.

    ahead: a spin whose block is done walks up to this many steps of the next
    block while the slower spins of its wavefront finish theirs, recording them
    in the ring (nz_walk_record_ring); the spins are re-sorted every
    sort_interval steps, after a block that ran nothing ahead."""

    def __init__(self, sub, x0, seed, g, u, dt, K=64, mode="mfma", overlap=True, ahead=None,
                 sort_interval=256):
        self.device = sub.device
        self.N = x0.shape[1]
        self.K = int(K)
        # K / 2 measured best (4e6 spins, 1024 waveforms, 1280 steps: K = 64
        # 0.527 s against 0.530 s at 64 steps and 0.537 s stopping at every
        # block end with a re-sort per block)
        self.ahead = self.K // 2 if ahead is None else int(ahead)
        if not 0 <= self.ahead <= self.K:
            raise ValueError("ahead must be in [0, K]")
        if mode == "mfma" and self.K not in (32, 64, 96, 128):
            raise ValueError("the MFMA contraction needs K = 32, 64, 96 or 128")
        # the MFMA contraction moves phase rows in 16-byte pieces: W is padded
        # to a multiple of 4 with zero waveforms (phases stay 0, not reported)
        self.W_out = g.shape[0]
        pad = (-g.shape[0]) % 4
        if pad:
            g = np.concatenate([g, np.zeros((pad, g.shape[1]), g.dtype)])
            u = np.concatenate([u, np.zeros((3, pad), u.dtype)], axis=1)
        self.W, self.T = g.shape
        self.mode = 0 if mode == "mfma" else 1
        self.seed, self.dt = int(seed), float(dt)
        d = self.device
        self.x0 = x0.to(d, torch.float32).contiguous()
        self.g = torch.from_numpy(np.ascontiguousarray(g, np.float32)).to(d)
        self.u = torch.from_numpy(np.ascontiguousarray(u, np.float32)).to(d)
        self.gsum = torch.from_numpy(g.astype(np.float64).sum(1).astype(np.float32)).to(d)
        # re-sort after whole blocks, about every sort_interval steps
        self.walker = rt.Walker(sub, self.N,
                                sort_interval=self.K * max(1, round(sort_interval / self.K)))
        self.phase = torch.empty((self.N, self.W), dtype=torch.float32, device=d)
        nb = 2 if overlap else 1
        # slot-major displacement record [N][3 ring] (a lane's steps
        # contiguous), a ring of blocks: block b at steps (b K) mod ring, the
        # steps run ahead after it; the block a contraction reads beside the
        # next walk stays clear of that walk's block and its run-ahead
        self.ring = (3 if overlap else 2) * self.K
        self.ld = 3 * self.ring
        self.rec = torch.zeros((self.N, self.ld), dtype=torch.float32, device=d)
        self.slot_spin = [torch.empty(self.N, dtype=torch.int32, device=d) for _ in range(nb)]
        self.bt_hi = [torch.empty((self.W, 3 * self.K), dtype=torch.int16, device=d)
                      for _ in range(nb)]
        self.bt_lo = [torch.empty_like(b) for b in self.bt_hi]
        self.bt_f32 = [torch.empty((self.W, 3 * self.K), dtype=torch.float32, device=d)
                       if self.mode == 1 else None for _ in range(nb)]
        self.signals_out = torch.empty((self.W, 2), dtype=torch.float64, device=d)
        self.walk_stream = torch.cuda.current_stream(d)
        self.gemm_stream = torch.cuda.Stream(d) if overlap else self.walk_stream
        if overlap:  # memory the contraction stream reads or writes outlives its queued work
            for t in ([self.phase, self.g, self.u, self.gsum, self.rec] + self.slot_spin +
                      self.bt_hi + self.bt_lo + [b for b in self.bt_f32 if b is not None]):
                t.record_stream(self.gemm_stream)
        self.walk_done = [torch.cuda.Event() for _ in range(nb)]
        self.gemm_done = [torch.cuda.Event() for _ in range(nb)]
        self._used = [False] * nb
        self._nb = nb
        self._idx = 0
        self.gemm_events = None  # list of (start, end) timing events when set to []

    def reset(self):
        L = lib()
        with torch.cuda.device(self.device), torch.cuda.stream(self.walk_stream):
            # contractions still in flight update the phases being reset
            self.walk_stream.wait_stream(self.gemm_stream)
            self.walker.set_positions(self.x0)
            check(L.nz_phase_init(_ptr(self.x0), self.N, _ptr(self.u), _ptr(self.gsum), self.W,
                                  _ptr(self.phase), _stream_ptr(self.device)), "nz_phase_init")
            e = torch.cuda.Event()
            e.record(self.walk_stream)
            self.gemm_stream.wait_event(e)
        self._idx = 0

    def block(self, t0, record_hook=None):
        """Walk and contract steps t0 .. t0 + K - 1 (fewer at the end)."""
        L = lib()
        k = min(self.K, self.T - t0)
        b = self._idx % self._nb
        pos0 = (self._idx * self.K) % self.ring
        self._idx += 1
        sa = ctypes.c_void_p(self.walk_stream.cuda_stream)
        sb = ctypes.c_void_p(self.gemm_stream.cuda_stream)
        with torch.cuda.device(self.device):
            if self._used[b]:  # block idx - 2's contraction still reads these buffers
                self.walk_stream.wait_event(self.gemm_done[b])
            ahead = min(self.ahead, self.T - t0 - k)  # never past the last step
            check(L.nz_walk_record_ring(self.walker.handle, self.seed & (2**64 - 1), k, ahead,
                                        self.dt, _ptr(self.rec), self.ld, self.ring, pos0,
                                        _ptr(self.slot_spin[b]), sa), "nz_walk_record_ring")
            if record_hook is not None:
                with torch.cuda.stream(self.walk_stream):
                    record_hook(t0, k, self.rec[:, 3 * pos0:3 * (pos0 + k)], self.slot_spin[b])
            self.walk_done[b].record(self.walk_stream)
            self.gemm_stream.wait_event(self.walk_done[b])
            if self.gemm_events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record(self.gemm_stream)
            check(L.nz_waveform_block(_ptr(self.g), self.T, _ptr(self.u), self.W, t0, self.K,
                                      _ptr(self.bt_hi[b]), _ptr(self.bt_lo[b]),
                                      _ptr(self.bt_f32[b]), sb), "nz_waveform_block")
            blk = ctypes.c_void_p(self.rec.data_ptr() + 12 * pos0)  # the block's first entry
            check(L.nz_phase_gemm(blk, self.ld, self.N, _ptr(self.slot_spin[b]),
                                  self.K, _ptr(self.bt_hi[b]), _ptr(self.bt_lo[b]),
                                  _ptr(self.bt_f32[b]), self.W, self.mode, _ptr(self.phase), sb),
                  "nz_phase_gemm")
            if self.gemm_events is not None:
                e1 = torch.cuda.Event(enable_timing=True)
                e1.record(self.gemm_stream)
                self.gemm_events.append((e0, e1))
            self.gemm_done[b].record(self.gemm_stream)
            self._used[b] = True

    def run(self, record_hook=None):
        self.reset()
        for t0 in range(0, self.T, self.K):
            self.block(t0, record_hook)

    def finish(self):
        """Make the current stream wait for every contraction issued so far."""
        e = torch.cuda.Event()
        e.record(self.gemm_stream)
        torch.cuda.current_stream(self.device).wait_event(e)

    def signals(self):
        """Complex [W] = sum exp(i phase) / N (FP64 accumulation)."""
        self.finish()
        with torch.cuda.device(self.device):
            check(lib().nz_phase_signals(_ptr(self.phase), self.N, self.W, _ptr(self.signals_out),
                                         _stream_ptr(self.device)), "nz_phase_signals")
        s = self.signals_out.cpu().numpy()[:self.W_out]
        return (s[:, 0] + 1j * s[:, 1]) / self.N

    def close(self):
        torch.cuda.synchronize(self.device)
        self.walker.close()
