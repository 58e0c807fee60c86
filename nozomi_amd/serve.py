"""Serving and batching forms of the walker.

``Server`` is a resident request loop on one GPU: simulation requests
(substrate, spin count, protocol) arrive at any time; substrates stay
resident in HBM and are reused across requests; a request is admitted when
its spins fit a contiguous free range of the ``Batch`` pool and its bytes fit
the HBM budget; admitted requests share every launch with those already
running (continuous batching: they join and leave at launch boundaries).  A
finished request's DWI signals are reduced in FP64 on the device and its
latency (arrival to result) recorded.

``run_serial`` runs the same requests back to back with one plain walker per
request: the baseline the batched/served forms are measured against.

A request's result does not depend on what else runs with it: every spin's
trajectory is keyed by (request seed, spin, step) and computed by the same
kernel code, so alone and in a mix it is bit-identical; only the FP64 signal
sums may differ in the last bits (order of the block partial sums).
"""
import collections
import dataclasses
import heapq
import threading
import time

import numpy as np
import torch

from . import runtime as rt
from .batch import Batch, BatchFull

# device bytes per spin slot held by a Batch (two state buffers, sort key,
# seeding buffer): 2 * 100 + 4 + 12
SLOT_BYTES = 216


@dataclasses.dataclass
class Request:
    """One simulation: `n` spins on substrate `key` (built by `make_geometry`
    when first seen), `seed` (Philox key), `dt` (ms) and the protocol's
    per-step coefficients `g`, `garea` (float32 [nsteps]) with `dirs` [3, W]
    and `scales` [W] of its signals.  `targets`: seeding structures (None =
    every structure and the box)."""
    rid: int
    key: object
    make_geometry: object
    n: int
    seed: int
    dt: float
    g: np.ndarray
    garea: np.ndarray
    dirs: np.ndarray
    scales: np.ndarray
    targets: object = None
    arrival: float = 0.0

    @property
    def nsteps(self):
        return int(self.g.size)

    @property
    def spin_steps(self):
        return self.n * self.nsteps


@dataclasses.dataclass
class Result:
    rid: int
    signals: np.ndarray      # complex [W] = sum exp(i phase) / n (None on error)
    arrival: float
    admitted: float
    finished: float
    error: str = None        # why the request failed (it never ran)

    @property
    def latency(self):
        return self.finished - self.arrival


class SubstrateCache:
    """Device substrates by key, built on first use, kept while referenced or
    while the byte budget allows (least recently used evicted first).
    ``prefetch`` builds them ahead (host geometry + device upload).  With
    tune=True the cell edge, list margin and steps per launch of a substrate
    come from the autotuner (per substrate family); ``params(key)`` returns
    them."""

    def __init__(self, device, budget_bytes, cell_size=0.0, margin=0.0, tune=False):
        self.device = device
        self.budget = budget_bytes
        self.cell_size, self.margin = cell_size, margin
        self.tune = tune
        self.items = collections.OrderedDict()   # key -> [sub, geom, bytes, refs, params]
        self.bytes = 0
        self.built = 0
        self.reused = 0

    def acquire(self, key, make_geometry):
        if key in self.items:
            self.items.move_to_end(key)
            self.items[key][3] += 1
            self.reused += 1
            return self.items[key][0], self.items[key][1]
        geom = make_geometry()
        params = {"cell": self.cell_size, "margin": self.margin, "k": None}
        if self.tune:
            from .autotune import tuned
            params = tuned(geom, device=self.device)
        sub = geom.device_substrate(cell_size=params["cell"], margin=params["margin"],
                                    device=self.device)
        nbytes = int(sub.info()["device_bytes"])
        self.items[key] = [sub, geom, nbytes, 1, params]
        self.bytes += nbytes
        self.built += 1
        self._evict()
        return sub, geom

    def params(self, key):
        """(cell, margin, k) the substrate was built with (k None: untuned)."""
        return self.items[key][4]

    def prefetch(self, requests):
        for r in requests:
            self.acquire(r.key, r.make_geometry)
            self.release(r.key)

    def pretune(self, requests):
        """Tunes every substrate family the requests use, ahead of serving
        (the autotuner keeps one choice per family for the process), so that
        no request waits for a tuning sweep.  Builds no device substrate.
        Returns {family key: (cell, margin, k)}."""
        from .autotune import family_key, tuned
        out, seen = {}, set()
        for r in requests:
            if r.key in seen:
                continue
            seen.add(r.key)
            geom = r.make_geometry()
            fk = family_key(geom)
            if fk not in out:
                p = tuned(geom, device=self.device)
                out[fk] = (p["cell"], p["margin"], p["k"])
        return out

    def release(self, key):
        if key in self.items:
            self.items[key][3] -= 1
            self._evict()

    def _evict(self):
        for k in list(self.items):
            if self.bytes <= self.budget:
                break
            if self.items[k][3] <= 0:
                self.bytes -= self.items[k][2]
                self.items[k][0].close()
                del self.items[k]


class Server:
    """Continuous-batching request loop on one GPU.

    capacity: spin slots of the pool; hbm_budget: bytes the server may hold
    for substrates and per-request protocol buffers beside the pool; k: steps
    per launch; launches: launches per scheduling round (2 replays the
    captured pair); tune: take cell edge, list margin and steps per launch
    from the autotuner (per substrate family) instead of the defaults and k."""

    def __init__(self, capacity, hbm_budget=8 << 30, max_sims=256, k=64, launches=2,
                 device=None, graph=True, cell_size=0.0, margin=0.0, cache=None, tune=False):
        self.device = rt._require_gpu(device)
        self.batch = Batch(capacity, max_sims=max_sims, device=self.device, graph=graph)
        self.k = int(k)
        self.launches = int(launches)
        self.hbm_budget = int(hbm_budget)
        self.subs = cache if cache is not None else SubstrateCache(
            self.device, self.hbm_budget // 2, cell_size, margin, tune=tune)
        self.queue = collections.deque()
        self.running = {}      # sim -> (request, admitted, tensors)
        self.results = {}
        self.lock = threading.Lock()
        self.proto_bytes = 0
        self.launched = 0

    # -- submission -----------------------------------------------------------
    def submit(self, req, now=None):
        if req.n > self.batch.capacity:
            raise ValueError(f"request {req.rid}: {req.n} spins exceed the pool ({self.batch.capacity})")
        req.arrival = time.perf_counter() if now is None else now
        with self.lock:
            self.queue.append(req)

    # -- scheduling -------------------------------------------------------------
    def _answer_error(self, req, e):
        """The request cannot run: drop it from the queue and answer it."""
        with self.lock:
            self.queue.popleft()
        now = time.perf_counter()
        self.results[req.rid] = Result(req.rid, None, req.arrival, now, now,
                                       error=f"{type(e).__name__}: {e}")

    def _admit(self):
        admitted = 0
        while True:
            with self.lock:
                if not self.queue:
                    return admitted
                req = self.queue[0]
            pb = 8 * req.nsteps + 16 * req.dirs.shape[1]
            if self.subs.bytes + self.proto_bytes + pb > self.hbm_budget and self.running:
                return admitted
            largest, _, free_sims = self.batch.free()
            if free_sims == 0 or largest < req.n:
                return admitted
            try:  # a substrate that cannot be built fails its request only
                sub, geom = self.subs.acquire(req.key, req.make_geometry)
            except Exception as e:
                self._answer_error(req, e)
                continue
            try:
                g = torch.from_numpy(np.ascontiguousarray(req.g, np.float32)).to(self.device)
                ga = torch.from_numpy(np.ascontiguousarray(req.garea, np.float32)).to(self.device)
                dirs = torch.from_numpy(np.ascontiguousarray(req.dirs, np.float32)).to(self.device)
                sc = torch.from_numpy(np.ascontiguousarray(req.scales, np.float32)).to(self.device)
                targets = req.targets if req.targets is not None else np.arange(geom.nstructures + 1)
                k = self.subs.params(req.key)["k"] or self.k
                sim = self.batch.add(sub, req.n, req.seed, req.nsteps, req.dt, g=g, garea=ga,
                                     k=k, targets=targets)
            except BatchFull:
                self.subs.release(req.key)
                return admitted
            except Exception as e:  # the request cannot run: answer it, keep serving
                self.subs.release(req.key)
                self._answer_error(req, e)
                continue
            with self.lock:
                self.queue.popleft()
            self.proto_bytes += pb
            self.running[sim] = (req, time.perf_counter(), (g, ga, dirs, sc), pb)
            admitted += 1

    def _harvest(self):
        done = []
        if not self.running:
            return done
        _, finished = self.batch.status_all()
        for sim, (req, t_adm, (g, ga, dirs, sc), pb) in list(self.running.items()):
            if finished[sim] == 1:
                s = self.batch.signals(sim, dirs, sc).cpu().numpy()
                now = time.perf_counter()
                self.results[req.rid] = Result(req.rid, (s[:, 0] + 1j * s[:, 1]) / req.n,
                                               req.arrival, t_adm, now)
                self.batch.remove(sim)
                self.subs.release(req.key)
                self.proto_bytes -= pb
                del self.running[sim]
                done.append(req.rid)
        return done

    def step(self):
        """One scheduling round: admit what fits, launch, collect finished
        requests.  Returns the ids finished in this round."""
        self._admit()
        if self.running:
            self.batch.run(self.launches)
            self.launched += self.launches
        return self._harvest()

    def idle(self):
        with self.lock:
            return not self.queue and not self.running

    def close(self):
        self.batch.close()


class _Progress:
    """Calls fn(done, total) at most every `every` seconds (long runs report
    as they go)."""

    def __init__(self, fn, total, every=20.0):
        self.fn, self.total, self.every = fn, total, every
        self.last = time.perf_counter()

    def __call__(self, done):
        if self.fn is not None and time.perf_counter() - self.last >= self.every:
            self.last = time.perf_counter()
            self.fn(done, self.total)


def serve_trace(server, requests, arrivals, poll=0.0, progress=None):
    """Drive `server` with requests arriving at the given times (seconds from
    the start, sorted).  Returns (results by id, makespan seconds).
    progress(done, total) is called every 20 s or so when given."""
    t0 = time.perf_counter()
    pending = list(zip(arrivals, range(len(requests))))
    heapq.heapify(pending)
    report = _Progress(progress, len(requests))
    while pending or not server.idle():
        report(len(server.results))
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            at, i = heapq.heappop(pending)
            server.submit(requests[i], now=t0 + at)
        if server.idle():
            if pending:
                time.sleep(max(0.0, min(pending[0][0] - now, 0.01)))
            continue
        server.step()
        if poll:
            time.sleep(poll)
    return server.results, time.perf_counter() - t0


def run_batched(requests, capacity, k=96, launches=2, device=None, graph=True, max_sims=256,
                cache=None, progress=None, tune=False):
    """Every request submitted at once, run to completion by one Server.
    Returns (results, seconds)."""
    srv = Server(capacity, hbm_budget=1 << 40, max_sims=max_sims, k=k, launches=launches,
                 device=device, graph=graph, cache=cache, tune=tune)
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in requests:
            srv.submit(r, now=t0)
        report = _Progress(progress, len(requests))
        while not srv.idle():
            srv.step()
            report(len(srv.results))
        torch.cuda.synchronize()
        return srv.results, time.perf_counter() - t0
    finally:
        srv.close()


def run_serial(requests, device=None, cache=None, progress=None):
    """Back-to-back plain launches: one walker per request, one request at a
    time (substrates cached the same way).  Returns (results, seconds); a
    result's arrival/finished are the start and end of its own run."""
    device = rt._require_gpu(device)
    cache = cache or SubstrateCache(device, 1 << 40)
    results = {}
    report = _Progress(progress, len(requests))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in requests:
        report(len(results))
        t_start = time.perf_counter()
        sub, geom = cache.acquire(r.key, r.make_geometry)
        w = rt.Walker(sub, r.n, sort_interval=cache.params(r.key)["k"])
        targets = r.targets if r.targets is not None else np.arange(geom.nstructures + 1)
        w.seed(targets=targets, seed=r.seed ^ 0x5EED5EED5EED)
        g = torch.from_numpy(np.ascontiguousarray(r.g, np.float32)).to(device)
        ga = torch.from_numpy(np.ascontiguousarray(r.garea, np.float32)).to(device)
        w.walk(r.seed, r.nsteps, r.dt, g=g, garea=ga, t2=False)
        mom = torch.empty((3, r.n), dtype=torch.float32, device=device)
        w.read(mom=mom)
        dirs = torch.from_numpy(np.ascontiguousarray(r.dirs, np.float32)).to(device)
        sc = torch.from_numpy(np.ascontiguousarray(r.scales, np.float32)).to(device)
        s = rt.signal_sums(mom, dirs, sc).cpu().numpy()
        now = time.perf_counter()
        results[r.rid] = Result(r.rid, (s[:, 0] + 1j * s[:, 1]) / r.n, t_start, t_start, now)
        cache.release(r.key)
        w.close()
    return results, time.perf_counter() - t0


def fifo_latencies(arrivals, durations):
    """Latencies of requests served one at a time in arrival order, each
    taking its measured duration."""
    t, out = 0.0, []
    for a, d in zip(arrivals, durations):
        t = max(t, a) + d
        out.append(t - a)
    return np.asarray(out)


def latency_stats(results):
    lat = np.array([r.latency for r in results.values()])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "n": int(lat.size)}
