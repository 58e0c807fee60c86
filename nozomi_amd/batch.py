"""Many simulations per launch (libnozomi_hip's nz_batch_*).

A ``Batch`` owns a pool of spin slots on one GPU.  Simulations of different
substrates, spin counts, lengths, time steps and protocols are added and
removed while others run; one launch advances every active simulation by up
to its ``k`` steps.  A simulation's spins follow exactly the trajectories a
``Walker`` gives them alone (same substrate, Philox key and start positions):
the random stream is keyed per spin and step, not per launch.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .runtime import _ptr, _require_gpu, _stream_ptr

NZ_EFULL = -2


class BatchFull(RuntimeError):
    """No free simulation entry or contiguous spin blocks (admission must wait)."""


class Batch:
    def __init__(self, capacity, max_sims=256, device=None, graph=True):
        self.device = _require_gpu(device)
        self.capacity = int(capacity)
        self.max_sims = int(max_sims)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().nz_batch_create(self.capacity, self.max_sims, ctypes.byref(h)),
                  "nz_batch_create")
        self.handle = h
        self._keep = {}
        if not graph:
            self.set_graph(False)

    def close(self):
        if getattr(self, "handle", None):
            with torch.cuda.device(self.device):
                lib().nz_batch_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, enable):
        check(lib().nz_batch_set_graph(self.handle, 1 if enable else 0), "nz_batch_set_graph")

    def add(self, sub, n, seed, nsteps, dt, g=None, garea=None, t2=False, k=64, pos=None,
            targets=None):
        """Add a simulation; returns its id.  sub: DeviceSubstrate; g/garea:
        float32 device tensors [nsteps] (kept alive until remove); pos:
        optional [3, n] start positions (device tensor), else seeding into
        `targets` (structure ids; None = every structure and the box)."""
        if (g is None) != (garea is None):
            raise ValueError("g and garea go together")
        if g is not None:
            if g.dtype != torch.float32 or garea.dtype != torch.float32:
                raise ValueError("g and garea must be float32 tensors")
            if g.numel() < nsteps or garea.numel() < nsteps:
                raise ValueError("g and garea need nsteps entries")
        pg, pa = _ptr(g, self.device, "g"), _ptr(garea, self.device, "garea")
        t = None
        if pos is None:
            if targets is None:
                targets = np.arange(sub.nstructures + 1)
            t = np.ascontiguousarray(targets, dtype=np.int32)
        else:
            pos = pos.to(self.device, torch.float32).contiguous()
            if pos.shape != (3, n):
                raise ValueError("pos must be [3, n]")
        sim = ctypes.c_int32(-1)
        with torch.cuda.device(self.device):
            rc = lib().nz_batch_add(self.handle, sub.handle, int(n), int(seed) & (2**64 - 1),
                                    int(nsteps), float(dt), pg, pa,
                                    _lib.NZ_WALK_T2 if t2 else 0, int(k),
                                    _ptr(pos, self.device, "pos"),
                                    None if t is None else t.ctypes.data_as(ctypes.c_void_p),
                                    0 if t is None else t.size, ctypes.byref(sim),
                                    _stream_ptr(self.device))
        if rc == NZ_EFULL:
            raise BatchFull(lib().nz_last_error_string().decode())
        check(rc, "nz_batch_add")
        self._keep[sim.value] = (sub, g, garea, pos)
        return sim.value

    def remove(self, sim):
        with torch.cuda.device(self.device):
            check(lib().nz_batch_remove(self.handle, int(sim)), "nz_batch_remove")
        self._keep.pop(sim, None)

    def run(self, launches=2):
        with torch.cuda.device(self.device):
            check(lib().nz_batch_run(self.handle, int(launches), _stream_ptr(self.device)),
                  "nz_batch_run")

    def status(self, sim):
        """(steps completed, done)."""
        step = ctypes.c_int64()
        done = ctypes.c_int32()
        with torch.cuda.device(self.device):
            check(lib().nz_batch_status(self.handle, int(sim), ctypes.byref(step), ctypes.byref(done)),
                  "nz_batch_status")
        return step.value, bool(done.value)

    def status_all(self):
        """(steps [max_sims], done [max_sims]); unused entries hold -1."""
        steps = np.zeros(self.max_sims, dtype=np.int64)
        done = np.zeros(self.max_sims, dtype=np.int32)
        with torch.cuda.device(self.device):
            check(lib().nz_batch_status_all(self.handle, steps.ctypes.data_as(ctypes.c_void_p),
                                            done.ctypes.data_as(ctypes.c_void_p)),
                  "nz_batch_status_all")
        return steps, done

    def free(self):
        """(largest contiguous free slots, free slots, free simulation entries)."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        check(lib().nz_batch_free(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              "nz_batch_free")
        return a.value, b.value, c.value

    def read(self, sim, n, pos=None, pos0=None, sig=None, mom=None):
        """Fill the given device tensors in spin order ([3, n] / [n])."""
        def out(t, name, numel):
            if t is not None and (t.dtype != torch.float32 or t.numel() < numel):
                raise ValueError(f"{name} must be a float32 tensor with at least {numel} entries")
            return _ptr(t, self.device, name)
        args = (out(pos, "pos", 3 * n), out(pos0, "pos0", 3 * n), out(sig, "sig", n),
                out(mom, "mom", 3 * n))
        with torch.cuda.device(self.device):
            check(lib().nz_batch_read(self.handle, int(sim), *args, _stream_ptr(self.device)),
                  "nz_batch_read")

    def signals(self, sim, dirs, scales=None, out=None):
        """FP64 [W, 2] (sum cos, sum sin) of scale_w u_w . m over the
        simulation's spins."""
        if dirs.dim() != 2 or dirs.shape[0] != 3 or dirs.dtype != torch.float32:
            raise ValueError("dirs must be a float32 tensor of shape (3, W)")
        W = dirs.shape[1]
        if scales is not None and (scales.dtype != torch.float32 or scales.numel() != W):
            raise ValueError("scales must be a float32 tensor of W entries")
        out = out if out is not None else torch.empty((W, 2), dtype=torch.float64,
                                                      device=self.device)
        if out.dtype != torch.float64 or out.numel() < 2 * W:
            raise ValueError("out must be a float64 tensor of 2*W entries")
        dirs = dirs.contiguous()
        args = (_ptr(dirs, self.device, "dirs"), _ptr(scales, self.device, "scales"),
                _ptr(out, self.device, "out"))
        with torch.cuda.device(self.device):
            check(lib().nz_batch_signals(self.handle, int(sim), int(W), *args,
                                         _stream_ptr(self.device)), "nz_batch_signals")
        return out
