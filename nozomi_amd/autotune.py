"""Autotuner of the walk's launch parameters for a substrate.

The walk kernel runs one spin per lane; what changes its speed on a given
substrate is the cell edge of the candidate grid (candidates per list against
re-bins), the list margin (list length against safety-box fallbacks) and the
steps per launch, which is also the spatial re-sort interval (sort cost
against wavefront coherence).  ``tune`` times a short walk of a subsample for
every combination and returns the fastest; results are cached per key.

``tuned`` is what the batched and served forms call for every substrate: it
tunes once per substrate family (packing density and sphere size, which set
the candidate counts and collision rates the choices trade against) and
reuses the choice for every substrate of the family.
"""
import itertools
import time

import numpy as np
import torch

from . import runtime as rt

_CACHE = {}


def tune(geom, key=None, n=65536, steps=512, dt=0.01, cells=(1.5, 1.75, 2.0, 2.25),
         margins=(0.8, 0.9), ks=(96, 128, 256), device=None, seed=1):
    """Returns {"cell": c, "margin": m, "k": k, "ns_per_spin_step": t,
    "table": [(t, c, m, k), ...]} for the geometry (a SimGeometry3D).  The
    timed walk (steps) spans at least two sort intervals of every k, so each
    candidate pays its re-sorts."""
    if key is not None and key in _CACHE:
        return _CACHE[key]
    device = rt._require_gpu(device)
    table = []
    for cell, margin in itertools.product(cells, margins):
        sub = geom.device_substrate(cell_size=cell, margin=margin, device=device)
        p = sub.seed(n, targets=np.arange(geom.nstructures + 1), seed=seed)
        for k in ks:
            w = rt.Walker(sub, n, sort_interval=k)
            w.set_positions(p)
            w.walk(seed, k, dt, t2=False)  # first sort and launch
            torch.cuda.synchronize(device)
            t0 = time.perf_counter()
            w.walk(seed, steps, dt, t2=False)
            torch.cuda.synchronize(device)
            table.append(((time.perf_counter() - t0) / (n * steps) * 1e9, cell, margin, k))
            w.close()
        sub.close()
    t, cell, margin, k = min(table)
    out = {"cell": cell, "margin": margin, "k": k, "ns_per_spin_step": t, "table": table}
    if key is not None:
        _CACHE[key] = out
    return out


def family_key(geom):
    """Substrates whose sphere density (spheres per um^3, in factors of 2) and
    mean sphere radius (in steps of 0.25 um) agree share their tuning.  Finer
    classes buy nothing: on the beaded-axon mixes the choices of neighbouring
    classes differ by less than the run-to-run spread of the timings."""
    sp = np.asarray(geom.spheres)
    if sp.shape[1] == 0:
        return ("empty",)
    vol = float(geom.Lx) * float(geom.Ly) * float(geom.Lz)
    dens = sp.shape[1] / vol
    return (int(round(np.log2(dens))), int(round(4 * float(sp[3].mean()))))


def tuned(geom, device=None, **kw):
    """The tuned (cell, margin, k) of the geometry's family (tunes on the first
    geometry of a family, then reuses it)."""
    return tune(geom, key=("family",) + family_key(geom), device=device, **kw)
