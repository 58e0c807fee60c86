"""ctypes wrapper of the FP64 CPU oracle (oracle/walker_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the nozomi_amd product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)
_up = ctypes.POINTER(ctypes.c_uint32)


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_version.restype = ctypes.c_int
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]
        L.orc_normals.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _dp]
        for f in (L.orc_step_inside, L.orc_step_outside):
            f.argtypes = [ctypes.c_double] * 7
            f.restype = ctypes.c_double
        L.orc_substrate_create.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double]
        L.orc_substrate_create.restype = ctypes.c_void_p
        L.orc_substrate_destroy.argtypes = [ctypes.c_void_p]
        L.orc_substrate_set_mode.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip]
        L.orc_walk.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_uint64,
                               ctypes.c_longlong, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp,
                               _dp, _lp, _lp, ctypes.c_int]
        L.orc_walk.restype = ctypes.c_int
        L.orc_seed.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_uint64,
                               ctypes.c_int, _dp, ctypes.c_int]
        L.orc_seed.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c, _up), _ptr(k, _up), _ptr(out, _up))
    return out


def normals(seed, stream, spin, step, draw0, n):
    out = np.zeros((n, 3), dtype=np.float64)
    lib().orc_normals(seed, stream, spin, step, draw0, n, _ptr(out))
    return out


def step_inside(dx, dy, dz, r, jx, jy, jz):
    return lib().orc_step_inside(dx, dy, dz, r, jx, jy, jz)


def step_outside(dx, dy, dz, r, jx, jy, jz):
    return lib().orc_step_outside(dx, dy, dz, r, jx, jy, jz)


class OracleSubstrate:
    """FP64 substrate: spheres [5,S] (cx,cy,cz,r,sid), D/T2 per structure with
    the box last, box lengths.  mode 0 = exact bins (complete candidate sets),
    mode 1 = the reference's padded segment lists."""

    def __init__(self, spheres, nstructures, D, T2, L, bin_size=1.0, halo=4.0):
        sp = np.ascontiguousarray(spheres, dtype=np.float64)
        if sp.ndim != 2 or sp.shape[0] != 5:
            raise ValueError("spheres must be [5, S]")
        self.S = sp.shape[1]
        self.nstructures = int(nstructures)
        self._D = np.ascontiguousarray(D, dtype=np.float64)
        self._T2 = np.ascontiguousarray(T2, dtype=np.float64)
        if self._D.size != self.nstructures + 1 or self._T2.size != self.nstructures + 1:
            raise ValueError("D and T2 need nstructures + 1 entries (box last)")
        Lx, Ly, Lz = (float(v) for v in L)
        self.L = (Lx, Ly, Lz)
        self._h = lib().orc_substrate_create(_ptr(sp), self.S, self.nstructures, _ptr(self._D),
                                             _ptr(self._T2), Lx, Ly, Lz, bin_size, halo)

    def set_reference_segments(self, nseg, seglist):
        nx, ny, nz = nseg
        sl = np.ascontiguousarray(seglist, dtype=np.int32)
        sps = sl.size // (nx * ny * nz)
        lib().orc_substrate_set_mode(self._h, 1, nx, ny, nz, sps, _ptr(sl, _ip))

    def set_exact(self):
        lib().orc_substrate_set_mode(self._h, 0, 1, 1, 1, 0, None)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_substrate_destroy(self._h)
            self._h = None

    def walk(self, pos, pos0, sig, seed, step0, nsteps, dt, g=None, garea=None, mom=None,
             stepmom=False, nsub=None, nthreads=0):
        """Advance the spins in place by nsteps.  Returns dict of stats."""
        N = pos.shape[1]
        for a in (pos, pos0):
            assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (3, N)
        assert sig.dtype == np.float64 and sig.size == N
        if g is not None:
            g = np.ascontiguousarray(g, dtype=np.float64)
            garea = np.ascontiguousarray(garea, dtype=np.float64)
            assert g.size >= nsteps and garea.size >= nsteps and mom is not None
        sm = np.zeros((nsteps, 6), dtype=np.float64) if stepmom else None
        stats = np.zeros(2, dtype=np.int64)
        lib().orc_walk(self._h, N, _ptr(pos), _ptr(pos0), _ptr(sig), int(seed), int(step0),
                       int(nsteps), float(dt), _ptr(g), _ptr(garea), _ptr(mom), _ptr(sm),
                       _ptr(nsub, _lp), _ptr(stats, _lp), int(nthreads))
        out = {"sub_steps": int(stats[0]), "capped": int(stats[1])}
        if stepmom:
            out["stepmom"] = sm
        return out

    def seed(self, N, targets, seed, max_attempts=100000, nthreads=0):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        pos = np.zeros((3, N), dtype=np.float64)
        failed = lib().orc_seed(self._h, N, _ptr(t, _ip), t.size, int(seed), max_attempts,
                                _ptr(pos), nthreads)
        if failed:
            raise RuntimeError(f"{failed} spins not seeded after {max_attempts} attempts")
        return pos
