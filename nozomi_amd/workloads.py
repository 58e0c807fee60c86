"""Seeded synthetic workloads of the batched and served forms.

``batching_mix``: independent simulations of uneven size — beaded-axon
substrates of 20-400 axons in 20-80 um boxes, log-uniform 1e4-1e6 spins,
2000-1e4 steps of 10 us, PGSE at 1-3 b-shells with 16-128 Fibonacci
directions each.  The axon count is drawn for an 80 um box and scaled by the
box area (a 20 um box cannot hold 400 axons of r0 ~ 1 um), so every substrate
has the packing density of the 200-axon / 60 um one.

``serving_mix``: a Poisson request stream; 70% small (1e4-5e4 spins), 25%
medium (2e5), 5% large (2e6 spins, 256 signals); each new request reuses an
earlier substrate with probability 0.3.

``scale_steps`` / ``scale_spins`` shrink the step and spin counts (for
bounded benchmark runs; 1.0 is the full configuration).
"""
import numpy as np

from .protocol import PGSEShells, fibonacci_hemisphere
from .serve import Request
from .substrate import beaded_geometry

DT = 0.01  # ms


def _geometry_factory(axons, L, seed):
    def make():
        return beaded_geometry(axons, L, seed, D=2.0, T2=100.0)
    return make


def _pgse_request(rid, key, make, n, T, ndir, nshell, seed, rng):
    te = T * DT
    delta, Delta = 0.25 * te, 0.75 * te
    shells = [1.0, 2.0, 3.0][:nshell]
    prot = PGSEShells(delta, Delta, DT, shells, fibonacci_hemisphere(ndir), te=te)
    return Request(rid=rid, key=key, make_geometry=make, n=int(n), seed=int(seed), dt=DT,
                   g=prot.g, garea=prot.garea, dirs=prot.dirs.astype(np.float32),
                   scales=prot.scales.astype(np.float32))


def _substrate_draw(rng, master):
    L = float(rng.uniform(20.0, 80.0))
    axons = max(4, int(round(rng.integers(20, 401) * (L / 80.0) ** 2)))
    sseed = int(rng.integers(1, 2**31 - 1)) ^ master
    return (axons, round(L, 3), sseed)


def batching_mix(nsims=256, seed=20261203, scale_steps=1.0, scale_spins=1.0):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(nsims):
        key = _substrate_draw(rng, seed)
        n = int(round(10 ** rng.uniform(4.0, 6.0) * scale_spins))
        T = max(8, int(round(rng.integers(2000, 10001) * scale_steps)))
        ndir = int(rng.integers(16, 129))
        nshell = int(rng.integers(1, 4))
        reqs.append(_pgse_request(i, key, _geometry_factory(*key), n, T, ndir, nshell,
                                  int(rng.integers(1, 2**62)), rng))
    return reqs


def serving_mix(rate=50.0, duration=120.0, seed=20261203, scale_steps=1.0, scale_spins=1.0,
                reuse=0.3):
    """Returns (requests, arrival times in seconds)."""
    rng = np.random.default_rng(seed)
    t, reqs, arrivals, keys = 0.0, [], [], []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        if keys and rng.random() < reuse:
            key = keys[int(rng.integers(len(keys)))]
        else:
            key = _substrate_draw(rng, seed)
            keys.append(key)
        u = rng.random()
        if u < 0.70:
            n, ndir, nshell = 10 ** rng.uniform(4.0, np.log10(5e4)), int(rng.integers(16, 129)), \
                int(rng.integers(1, 4))
        elif u < 0.95:
            n, ndir, nshell = 2e5, int(rng.integers(16, 129)), int(rng.integers(1, 4))
        else:
            n, ndir, nshell = 2e6, 128, 2   # 256 signals
        T = max(8, int(round(rng.integers(2000, 10001) * scale_steps)))
        reqs.append(_pgse_request(len(reqs), key, _geometry_factory(*key),
                                  max(256, int(round(n * scale_spins))), T, ndir, nshell,
                                  int(rng.integers(1, 2**62)), rng))
        arrivals.append(t)
    return reqs, np.asarray(arrivals)


def pregenerate(requests):
    """Generate every distinct substrate geometry once (host side: the
    reference's substrate generator stands in for this step) and make the
    requests return it; device substrates are still built on first use."""
    geoms = {}
    for r in requests:
        if r.key not in geoms:
            geoms[r.key] = r.make_geometry()
    for r in requests:
        g = geoms[r.key]
        r.make_geometry = (lambda g=g: g)
    return geoms


def total_spin_steps(requests):
    return float(sum(r.spin_steps for r in requests))
