"""Device-side objects of the walker: substrate handle, resident spin state and
a GPUArray-like view.

Torch (ROCm) only provides device memory and streams here; all arithmetic on
the walk path runs in libnozomi_hip.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import NzSubstrateInfo, check, lib

_c_vp = ctypes.c_void_p

# rejection-seeding attempts per spin (as kSeedAttempts in the library: a
# target filling 0.2% of the box needs ~500 on average; attempts are keyed
# per spin, so the bound changes no spin that fewer attempts placed)
SEED_ATTEMPTS = 1 << 20


def _require_gpu(device):
    if not torch.cuda.is_available():
        raise _lib.NozomiError("nozomi_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device(device if device is not None else "cuda:0")


def _stream_ptr(device):
    return _c_vp(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t, device=None, name="tensor"):
    """Device address of a tensor handed to a kernel (NULL for None).

    The kernels dereference it as a dense array on the walker's GPU, so a
    host tensor, a tensor on another GPU or a strided view is refused here,
    before any launch: a host address would fault the GPU, and a view would be
    read with the wrong layout."""
    if t is None:
        return _c_vp(0)
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch.Tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor (got one on {t.device})")
    if device is not None and t.device != torch.device(device):
        raise ValueError(f"{name} is on {t.device}, the walker on {torch.device(device)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous (call .contiguous())")
    return _c_vp(t.data_ptr())


class DeviceArray:
    """Thin stand-in for pycuda.gpuarray.GPUArray over a torch tensor.

    The reference's tests read device results as ``spins_d[0,:].get()`` and
    ``((a - b)**2).get()`` (tests/validation/narrow_pulse/test_01...py:40-52);
    this class offers .get(), slicing, elementwise arithmetic and .shape."""

    __array_priority__ = 1000

    def __init__(self, tensor):
        self.tensor = tensor

    def get(self):
        torch.cuda.synchronize(self.tensor.device)
        return self.tensor.detach().cpu().numpy()

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return np.dtype(str(self.tensor.dtype).replace("torch.", ""))

    @property
    def size(self):
        return self.tensor.numel()

    def __len__(self):
        return self.tensor.shape[0]

    def __getitem__(self, idx):
        return DeviceArray(self.tensor[idx])

    def copy(self):
        return DeviceArray(self.tensor.clone())

    @staticmethod
    def _v(o):
        return o.tensor if isinstance(o, DeviceArray) else o

    def __add__(self, o): return DeviceArray(self.tensor + self._v(o))
    def __radd__(self, o): return DeviceArray(self._v(o) + self.tensor)
    def __sub__(self, o): return DeviceArray(self.tensor - self._v(o))
    def __rsub__(self, o): return DeviceArray(self._v(o) - self.tensor)
    def __mul__(self, o): return DeviceArray(self.tensor * self._v(o))
    def __rmul__(self, o): return DeviceArray(self._v(o) * self.tensor)
    def __truediv__(self, o): return DeviceArray(self.tensor / self._v(o))
    def __rtruediv__(self, o): return DeviceArray(self._v(o) / self.tensor)
    def __pow__(self, o): return DeviceArray(self.tensor ** self._v(o))
    def __neg__(self): return DeviceArray(-self.tensor)

    def __repr__(self):
        return f"DeviceArray({self.shape}, {self.tensor.dtype}, {self.tensor.device})"


class DeviceSubstrate:
    """Spheres + cell grid resident on the GPU (nz_substrate_create).

    spheres: [5, S] (cx, cy, cz, r, sid) float64; D, T2: nstructures + 1
    values with the box last; L: (Lx, Ly, Lz)."""

    def __init__(self, spheres, nstructures, D, T2, L, cell_size=0.0, margin=0.0, device=None):
        self.device = _require_gpu(device)
        sp = np.ascontiguousarray(spheres, dtype=np.float64)
        if sp.ndim != 2 or sp.shape[0] != 5:
            raise ValueError("spheres must be a [5, S] array")
        D = np.ascontiguousarray(D, dtype=np.float64)
        T2 = np.ascontiguousarray(T2, dtype=np.float64)
        if D.size != nstructures + 1 or T2.size != nstructures + 1:
            raise ValueError("D and T2 need nstructures + 1 values (box last)")
        self.nstructures = int(nstructures)
        self.L = tuple(float(v) for v in L)
        self.spheres = sp
        h = _c_vp()
        with torch.cuda.device(self.device):
            check(lib().nz_substrate_create(sp.ctypes.data_as(_c_vp), sp.shape[1], self.nstructures,
                                            D.ctypes.data_as(_c_vp), T2.ctypes.data_as(_c_vp),
                                            self.L[0], self.L[1], self.L[2], float(cell_size),
                                            float(margin), ctypes.byref(h)), "nz_substrate_create")
        self.handle = h

    def info(self):
        inf = NzSubstrateInfo()
        check(lib().nz_substrate_info(self.handle, ctypes.byref(inf)), "nz_substrate_info")
        return {
            "nspheres": inf.nspheres, "nstructures": inf.nstructures,
            "ncell": tuple(inf.ncell), "cell": tuple(inf.cell), "margin": inf.margin,
            "ncand_total": inf.ncand_total, "ncand_max": inf.ncand_max,
            "ncand_mean_box": inf.ncand_mean_box, "device_bytes": inf.device_bytes,
            "ncand_records": inf.ncand_records,
        }

    def grid(self):
        """The device-built cell grid as numpy arrays (nz_substrate_export_grid):
        cell_start int32 [nloc^3 + 1], cell_rank int32 [ncell^3], cand float32
        [2 ncand_records, 4] records (padding records at the end)."""
        inf = self.info()
        nloc = [n + 2 for n in inf["ncell"]]
        start = np.empty(nloc[0] * nloc[1] * nloc[2] + 1, np.int32)
        rank = np.empty(int(np.prod(inf["ncell"])), np.int32)
        cand = np.empty((2 * inf["ncand_records"], 4), np.float32)
        check(lib().nz_substrate_export_grid(self.handle, start.ctypes.data_as(_c_vp),
                                             rank.ctypes.data_as(_c_vp),
                                             cand.ctypes.data_as(_c_vp)),
              "nz_substrate_export_grid")
        return start, rank, cand

    def reach(self):
        """The device-built reach entries (nz_substrate_export_reach):
        reach_start int32 [nloc^3], reach uint32 [nentries, 2]."""
        inf = self.info()
        nloc = [n + 2 for n in inf["ncell"]]
        n = ctypes.c_int64(0)
        check(lib().nz_substrate_export_reach(self.handle, ctypes.byref(n), None, None),
              "nz_substrate_export_reach")
        rs = np.empty(nloc[0] * nloc[1] * nloc[2], np.int32)
        reach = np.empty((n.value, 2), np.uint32)
        check(lib().nz_substrate_export_reach(self.handle, None, rs.ctypes.data_as(_c_vp),
                                              reach.ctypes.data_as(_c_vp)),
              "nz_substrate_export_reach")
        return rs, reach

    def close(self):
        if getattr(self, "handle", None):
            lib().nz_substrate_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed(self, n, targets=None, seed=0, max_attempts=SEED_ATTEMPTS, out=None):
        """Rejection-seed n spins into the target structures; returns [3, n]."""
        pos = out if out is not None else torch.empty((3, n), dtype=torch.float32, device=self.device)
        if pos.dtype != torch.float32 or pos.numel() < 3 * n:
            raise ValueError(f"out must be a float32 tensor with 3*{n} entries")
        pp = _ptr(pos, self.device, "out")
        failed = torch.zeros(1, dtype=torch.int32, device=self.device)
        t = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
        with torch.cuda.device(self.device):
            check(lib().nz_seed(self.handle, n, None if t is None else t.ctypes.data_as(_c_vp),
                                0 if t is None else t.size, int(seed) & (2**64 - 1),
                                int(max_attempts), pp, _ptr(failed),
                                _stream_ptr(self.device)), "nz_seed")
        nf = int(failed.item())
        if nf:
            raise RuntimeError(f"{nf} spins could not be seeded in {max_attempts} attempts: "
                               "the target structures are empty or too small")
        return pos


class Walker:
    """Resident state of n spins on a substrate (nz_walker_*).

    The library keeps the state cell-ordered in double precision; read()
    returns the reference's views in spin order."""

    def __init__(self, sub, n, counters=False, sort_interval=None):
        self.sub = sub
        self.device = sub.device
        self.n = int(n)
        h = _c_vp()
        with torch.cuda.device(self.device):
            check(lib().nz_walker_create(sub.handle, self.n,
                                         _lib.NZ_WALKER_COUNTERS if counters else 0,
                                         ctypes.byref(h)), "nz_walker_create")
        self.handle = h
        self.counters_enabled = bool(counters)
        if sort_interval is not None:
            self.set_sort_interval(sort_interval)

    def close(self):
        if getattr(self, "handle", None):
            lib().nz_walker_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sort_interval(self, steps):
        check(lib().nz_walker_set_sort_interval(self.handle, int(steps)), "nz_walker_set_sort_interval")

    @property
    def step(self):
        v = ctypes.c_int64()
        check(lib().nz_walker_get_step(self.handle, ctypes.byref(v)), "nz_walker_get_step")
        return v.value

    @step.setter
    def step(self, value):
        check(lib().nz_walker_set_step(self.handle, int(value)), "nz_walker_set_step")

    def set_positions(self, pos):
        """Start positions [3, n] float32 (tensor on the walker's device)."""
        if tuple(pos.shape) != (3, self.n) or pos.dtype != torch.float32:
            raise ValueError(f"positions must be a float32 tensor of shape (3, {self.n})")
        pos = pos.to(self.device).contiguous()
        with torch.cuda.device(self.device):
            check(lib().nz_walker_set_positions(self.handle, _ptr(pos), _stream_ptr(self.device)),
                  "nz_walker_set_positions")
        self._keep = pos

    def seed(self, targets=None, seed=0, max_attempts=SEED_ATTEMPTS):
        failed = torch.zeros(1, dtype=torch.int32, device=self.device)
        t = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
        with torch.cuda.device(self.device):
            check(lib().nz_walker_seed(self.handle, None if t is None else t.ctypes.data_as(_c_vp),
                                       0 if t is None else t.size, int(seed) & (2**64 - 1),
                                       int(max_attempts), _ptr(failed), _stream_ptr(self.device)),
                  "nz_walker_seed")
        nf = int(failed.item())
        if nf:
            raise RuntimeError(f"{nf} spins could not be seeded in {max_attempts} attempts: "
                               "the target structures are empty or too small")

    def walk(self, seed, nsteps, dt, g=None, garea=None, stepmom=None, t2=True):
        """Advance nsteps (nz_walk).  g/garea: float32 device tensors [nsteps];
        stepmom: float64 device tensor [nsteps, 6] (sums are added)."""
        if (g is None) != (garea is None):
            raise ValueError("g and garea go together")
        if g is not None:
            if g.dtype != torch.float32 or garea.dtype != torch.float32:
                raise ValueError("g and garea must be float32 tensors")
            if g.numel() < nsteps or garea.numel() < nsteps:
                raise ValueError("g and garea need nsteps entries")
        if stepmom is not None and (stepmom.dtype != torch.float64 or stepmom.numel() < 6 * nsteps):
            raise ValueError("stepmom must be a float64 tensor with 6*nsteps entries")
        dev = self.device
        pg, pa = _ptr(g, dev, "g"), _ptr(garea, dev, "garea")
        pm = _ptr(stepmom, dev, "stepmom")
        with torch.cuda.device(dev):
            check(lib().nz_walk(self.handle, int(seed) & (2**64 - 1), int(nsteps), float(dt),
                                pg, pa, pm, _lib.NZ_WALK_T2 if t2 else 0, _stream_ptr(dev)),
                  "nz_walk")

    def _out(self, t, name, dtype, numel):
        if t is None:
            return _ptr(None)
        if t.dtype != dtype or t.numel() < numel:
            raise ValueError(f"{name} must be a {dtype} tensor with at least {numel} entries")
        return _ptr(t, self.device, name)

    def read(self, pos=None, pos0=None, sig=None, mom=None):
        """Fill the given tensors (spin order): pos/pos0/mom [3, n], sig [n]."""
        f32, n = torch.float32, self.n
        args = (self._out(pos, "pos", f32, 3 * n), self._out(pos0, "pos0", f32, 3 * n),
                self._out(sig, "sig", f32, n), self._out(mom, "mom", f32, 3 * n))
        with torch.cuda.device(self.device):
            check(lib().nz_walker_read(self.handle, *args, _stream_ptr(self.device)),
                  "nz_walker_read")

    def read_f64(self):
        """FP64 wrapped positions and unwrapped displacements, numpy [3, n]."""
        pos = torch.empty((3, self.n), dtype=torch.float64, device=self.device)
        disp = torch.empty((3, self.n), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(lib().nz_walker_read_f64(self.handle, _ptr(pos), _ptr(disp),
                                           _stream_ptr(self.device)), "nz_walker_read_f64")
        return pos.cpu().numpy(), disp.cpu().numpy()

    def counters(self):
        """Per-spin (scatter sub-steps, FP64 scans), numpy int32 [n] each."""
        a = torch.empty(self.n, dtype=torch.int32, device=self.device)
        b = torch.empty(self.n, dtype=torch.int32, device=self.device)
        with torch.cuda.device(self.device):
            check(lib().nz_walker_counters(self.handle, _ptr(a), _ptr(b), _stream_ptr(self.device)),
                  "nz_walker_counters")
        return a.cpu().numpy(), b.cpu().numpy()

    def time_kernels(self, enable=True):
        """Bracket every walk-kernel launch with HIP events on its stream."""
        check(lib().nz_walker_time_kernels(self.handle, 1 if enable else 0), "nz_walker_time_kernels")

    def kernel_time(self):
        """(total ms, launches) of the walk-kernel launches timed since the
        last call; waits for them."""
        ms = ctypes.c_double()
        k = ctypes.c_int64()
        with torch.cuda.device(self.device):
            check(lib().nz_walker_kernel_time(self.handle, ctypes.byref(ms), ctypes.byref(k)),
                  "nz_walker_kernel_time")
        return ms.value, k.value

    def displacement_moments(self, out=None):
        out = out if out is not None else torch.empty(6, dtype=torch.float64, device=self.device)
        po = self._out(out, "out", torch.float64, 6)
        with torch.cuda.device(self.device):
            check(lib().nz_displacement_moments(self.handle, po, _stream_ptr(self.device)),
                  "nz_displacement_moments")
        return out


def _projection_args(mom, dirs, scale):
    if mom.dim() != 2 or mom.shape[0] != 3 or mom.dtype != torch.float32:
        raise ValueError("mom must be a float32 tensor of shape (3, n)")
    if dirs.dim() != 2 or dirs.shape[0] != 3 or dirs.dtype != torch.float32:
        raise ValueError("dirs must be a float32 tensor of shape (3, W)")
    W = dirs.shape[1]
    if scale is not None and (scale.dtype != torch.float32 or scale.numel() != W):
        raise ValueError("scale must be a float32 tensor of W entries")
    dev = mom.device
    # one dense copy: its address goes to the kernel and the caller holds it
    # until the launch is queued (a second .contiguous() would be another copy)
    dc = dirs.contiguous()
    return (_ptr(mom, dev, "mom"), _ptr(dc, dev, "dirs"), _ptr(scale, dev, "scale"), dc)


def project_phase(mom, dirs, scale=None, out=None):
    """phase[w, i] = scale[w] * dirs[:, w] . mom[:, i]  (nz_project_phase)."""
    n = mom.shape[1]
    W = dirs.shape[1]
    pm, pd, ps, keep = _projection_args(mom, dirs, scale)
    out = out if out is not None else torch.empty((W, n), dtype=torch.float32, device=mom.device)
    if out.dtype != torch.float32 or out.numel() < W * n:
        raise ValueError("out must be a float32 tensor of W*n entries")
    check(lib().nz_project_phase(n, W, pm, pd, ps, _ptr(out, mom.device, "out"),
                                 _stream_ptr(mom.device)), "nz_project_phase")
    del keep
    return out


def signal_sums(mom, dirs, scale=None, out=None):
    """FP64 [W, 2] sums of (cos, sin) of the projected phases (nz_signal_reduce)."""
    n = mom.shape[1]
    W = dirs.shape[1]
    pm, pd, ps, keep = _projection_args(mom, dirs, scale)
    out = out if out is not None else torch.empty((W, 2), dtype=torch.float64, device=mom.device)
    if out.dtype != torch.float64 or out.numel() < 2 * W:
        raise ValueError("out must be a float64 tensor of 2*W entries")
    check(lib().nz_signal_reduce(n, W, pm, pd, ps, _ptr(out, mom.device, "out"),
                                 _stream_ptr(mom.device)), "nz_signal_reduce")
    del keep
    return out
