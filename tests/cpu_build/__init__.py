"""CPU build of the walker's per-spin core (nozomi_amd/csrc/nz_walk_core.h +
nz_grid_build.h), compiled with hipcc as host code.  Test infrastructure: lets
the CPU suite check the kernel logic against the FP64 oracle without a GPU."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "nz_walk_cpu.cpp")
LIB = os.path.join(HERE, "libnz_walk_cpu.so")
DEPS = [SRC] + [os.path.join(ROOT, "nozomi_amd", "csrc", f) for f in
                ("nz_walk_core.h", "nz_grid_build.h", "nz_internal.h", "nz_rng.h")]
_vp = ctypes.c_void_p
_lib = None


def build():
    if os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(d) for d in DEPS):
        return
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
                           "-o", LIB, SRC], stderr=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB)
        L.nzcpu_create.restype = _vp
        L.nzcpu_create.argtypes = [_vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.nzcpu_destroy.argtypes = [_vp]
        L.nzcpu_init.argtypes = [_vp, ctypes.c_int, _vp]
        L.nzcpu_walk.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                 _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
        L.nzcpu_state_f64.argtypes = [_vp, _vp, _vp]
        L.nzcpu_info.argtypes = [_vp, _vp]
        L.nzcpu_grid.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.nzcpu_scan_compare.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
        L.nzcpu_reach_check.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.nzcpu_reach.argtypes = [_vp, _vp, _vp, _vp]
        L.nzcpu_reach_params.argtypes = [_vp, _vp]
        L.nzcpu_normals.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_int, _vp]
        _lib = L
    return _lib


class CpuWalker:
    """The walk kernel's per-lane logic run on the host, one spin at a time."""

    def __init__(self, geom, cell=2.0, margin=1.0):
        sp = np.ascontiguousarray(geom.spheres, dtype=np.float64)
        self._keep = (sp, np.ascontiguousarray(geom.structure_D(), np.float64),
                      np.ascontiguousarray(geom.structure_T2(), np.float64))
        self.L = geom.Lx
        self.h = lib().nzcpu_create(sp.ctypes.data, sp.shape[1], geom.nstructures,
                                    self._keep[1].ctypes.data, self._keep[2].ctypes.data,
                                    geom.Lx, geom.Ly, geom.Lz, cell, margin)
        if not self.h:
            raise ValueError("grid build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().nzcpu_destroy(self.h)

    def scan_compare(self, pos, jump, jr, sid):
        """Per probe, which keys of the two-phase scan differ from visiting
        every candidate: int32 [n, 4] (kin, kout, sid_new, guards beyond the
        allowed difference)."""
        n = pos.shape[1]
        pos = np.ascontiguousarray(pos, np.float64)
        jump = np.ascontiguousarray(jump, np.float64)
        jr = np.ascontiguousarray(jr, np.float64)
        sid = np.ascontiguousarray(sid, np.float32)
        out = np.zeros((n, 4), np.int32)
        lib().nzcpu_scan_compare(self.h, n, pos.ctypes.data, jump.ctypes.data, jr.ctypes.data,
                                 sid.ctypes.data, out.ctypes.data)
        return out

    def reach_check(self, pos, jump, jr):
        """Per probe: int32 [n, 3] = (candidates the segment touches that the
        FP32 reach test rejects, candidates it touches (FP64), candidates
        the reach test passes)."""
        n = pos.shape[1]
        pos = np.ascontiguousarray(pos, np.float64)
        jump = np.ascontiguousarray(jump, np.float64)
        jr = np.ascontiguousarray(jr, np.float64)
        out = np.zeros((n, 3), np.int32)
        lib().nzcpu_reach_check(self.h, n, pos.ctypes.data, jump.ctypes.data, jr.ctypes.data,
                                out.ctypes.data)
        return out

    def reach(self):
        """The host builder's (reach_start int32 [nloc^3], reach uint32
        [entries, 2])."""
        sizes = np.zeros(2, np.int64)
        lib().nzcpu_reach(self.h, None, None, sizes.ctypes.data)
        rs = np.empty(sizes[1], np.int32)
        reach = np.empty((sizes[0], 2), np.uint32)
        lib().nzcpu_reach(self.h, rs.ctypes.data, reach.ctypes.data, sizes.ctypes.data)
        return rs, reach

    def reach_params(self):
        """(lambda, sigma, sigma (1 + 1/lambda), eps) of the reach test."""
        out = np.zeros(4, np.float64)
        lib().nzcpu_reach_params(self.h, out.ctypes.data)
        return tuple(out)

    def grid(self):
        """The host builder's (cell_start, cell_rank, cand [2 records, 4]), padding
        records included."""
        sizes = np.zeros(3, np.int64)
        lib().nzcpu_grid(self.h, None, None, None, sizes.ctypes.data)
        start = np.empty(sizes[0], np.int32)
        rank = np.empty(sizes[1], np.int32)
        cand = np.empty((sizes[2] // 4, 4), np.float32)
        lib().nzcpu_grid(self.h, start.ctypes.data, rank.ctypes.data, cand.ctypes.data,
                         sizes.ctypes.data)
        return start, rank, cand

    def init(self, pos):
        self.pos_init = np.ascontiguousarray(pos, dtype=np.float32)
        self.N = self.pos_init.shape[1]
        lib().nzcpu_init(self.h, self.N, self.pos_init.ctypes.data)

    def walk(self, seed, step0, nsteps, dt, g=None, garea=None, t2=True):
        N = self.N
        out = {k: np.zeros((3, N), np.float32) for k in ("pos", "pos0", "mom")}
        nsub = np.zeros(N, np.int32)
        nguard = np.zeros(N, np.int32)
        logsig = np.zeros(N, np.float32)
        gg = None if g is None else np.ascontiguousarray(g, np.float32)
        ga = None if g is None else np.ascontiguousarray(garea, np.float32)
        lib().nzcpu_walk(self.h, seed, step0, nsteps, dt,
                         None if gg is None else gg.ctypes.data, None if ga is None else ga.ctypes.data,
                         int(t2), out["pos"].ctypes.data, out["pos0"].ctypes.data, nsub.ctypes.data,
                         out["mom"].ctypes.data if g is not None else None, logsig.ctypes.data,
                         nguard.ctypes.data)
        pd = np.zeros((3, N))
        dd = np.zeros((3, N))
        lib().nzcpu_state_f64(self.h, pd.ctypes.data, dd.ctypes.data)
        out.update(nsub=nsub, nguard=nguard, logsig=logsig, pos_f64=pd, disp_f64=dd)
        return out
