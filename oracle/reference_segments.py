"""Numpy restatement of the reference's per-segment candidate lists.

TEST INFRASTRUCTURE ONLY (feeds the oracle's candidate mode 1).

Follows DiffSim3d.set_segments (simulation_toolkit/simulation_engine/
diffsim3d.py:107-174): a sphere belongs to segment (n,m,p) when the largest of
its six signed AABB-to-face distances is below jump_tol = 1 um (:110,128-141);
every list is padded to the largest count with the next-nearest spheres via
np.argsort, then sorted by sphere index (:167).  The same numpy calls are used
so that the padded extras come out in the same order.
"""
import numpy as np


def reference_segments(spheres, L, nseg, jump_tol=1.0):
    """spheres: [5,S] (cx,cy,cz,r,sid); L: (Lx,Ly,Lz); nseg: (nx,ny,nz).

    Returns (seglist int32 [nx*ny*nz*sps], sps)."""
    nsegx, nsegy, nsegz = nseg
    Lx, Ly, Lz = L
    dLx, dLy, dLz = Lx / nsegx, Ly / nsegy, Lz / nsegz
    sp = np.asarray(spheres, dtype=np.float64)

    def in_segment(n, m, p):
        inx_min = n * dLx - sp[3, :] - sp[0, :] - Lx / 2
        inx_max = sp[0, :] - ((n + 1) * dLx + sp[3, :] - Lx / 2)
        iny_min = m * dLy - sp[3, :] - sp[1, :] - Ly / 2
        iny_max = sp[1, :] - ((m + 1) * dLy + sp[3, :] - Ly / 2)
        inz_min = p * dLz - sp[3, :] - sp[2, :] - Lz / 2
        inz_max = sp[2, :] - ((p + 1) * dLz + sp[3, :] - Lz / 2)
        v = np.maximum(inx_min, inx_max)
        v = np.maximum(v, iny_min)
        v = np.maximum(v, iny_max)
        v = np.maximum(v, inz_min)
        return np.maximum(v, inz_max)

    sps = 0
    for n in np.arange(nsegx):
        for m in np.arange(nsegy):
            for p in np.arange(nsegz):
                sps = max(sps, int(np.count_nonzero(in_segment(n, m, p) < jump_tol)))
    seg = np.zeros([nsegx * nsegy * nsegz * sps], dtype=np.intc)
    for n in np.arange(nsegx):
        for m in np.arange(nsegy):
            for p in np.arange(nsegz):
                idx = sps * (p * nsegx * nsegy + m * nsegx + n)
                seg[idx:idx + sps] = np.sort(np.argsort(in_segment(n, m, p))[:sps])
    return seg.astype(np.int32), sps
