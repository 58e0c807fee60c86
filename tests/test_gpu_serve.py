"""The served form: requests that arrive while others run are admitted as
pool space frees up, share the launches, and each result equals the result of
the same request run alone with a plain walker (within the FP32/FP64
summation-order tolerance of the signal sums).  Also the autotuner's choice."""
import numpy as np
import pytest

from nozomi_amd.autotune import tune
from nozomi_amd.serve import Server, latency_stats, run_batched, run_serial, serve_trace
from nozomi_amd.substrate import beaded_geometry
from nozomi_amd.workloads import batching_mix, serving_mix

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-6  # |signal difference| (normalised by N), same moments summed in another order


def _compare(a, b):
    assert set(a) == set(b)
    worst = max(np.abs(a[r].signals - b[r].signals).max() for r in a)
    print(f"max |served - alone| = {worst:.2e}")
    assert worst < TOL


def test_served_results_equal_alone():
    reqs, arrivals = serving_mix(rate=40.0, duration=0.5, seed=5, scale_steps=0.05,
                                 scale_spins=0.05)
    assert len(reqs) >= 8
    # a pool smaller than the offered spins: later arrivals wait for space
    cap = max(r.n for r in reqs) + 2048
    srv = Server(cap, max_sims=8, k=32, device=DEV)
    got, makespan = serve_trace(srv, reqs, arrivals)
    print(f"{len(got)} requests in {makespan:.2f} s, latency {latency_stats(got)}; "
          f"substrates built {srv.subs.built}, reused {srv.subs.reused}")
    srv.close()
    alone, _ = run_serial(reqs, device=DEV)
    _compare(got, alone)


def test_failing_request_is_answered_and_serving_continues():
    """A request whose substrate cannot be built (its geometry callback
    raises) and one that cannot be seeded are answered with an error; the
    requests around them still run and equal their results alone."""
    import dataclasses
    reqs, _ = serving_mix(rate=40.0, duration=0.3, seed=7, scale_steps=0.05, scale_spins=0.05)
    reqs = reqs[:4]

    def broken_geometry():
        raise OSError("substrate file unreadable")

    bad_sub = dataclasses.replace(reqs[1], rid=1001, key=("broken", 1),
                                  make_geometry=broken_geometry)
    bad_seed = dataclasses.replace(reqs[2], rid=1002, targets=np.array([10**6]))
    cap = sum(r.n for r in reqs) + 2048
    srv = Server(cap, max_sims=8, k=32, device=DEV)
    got, _ = serve_trace(srv, [reqs[0], bad_sub, reqs[1], bad_seed, reqs[2], reqs[3]],
                         [0.0] * 6)
    srv.close()
    assert got[1001].signals is None and "substrate file unreadable" in got[1001].error
    assert got[1002].signals is None and got[1002].error
    good = {r.rid: got[r.rid] for r in reqs}
    assert all(v.error is None for v in good.values())
    alone, _ = run_serial(reqs, device=DEV)
    _compare(good, alone)


def test_batched_mix_equals_alone():
    reqs = batching_mix(nsims=12, seed=9, scale_steps=0.03, scale_spins=0.02)
    got, secs = run_batched(reqs, capacity=sum(r.n for r in reqs) + 256 * len(reqs), k=32,
                            device=DEV)
    alone, secs_serial = run_serial(reqs, device=DEV)
    print(f"batched {secs:.2f} s, serial {secs_serial:.2f} s")
    _compare(got, alone)


def test_autotune_picks_a_measured_option():
    geom = beaded_geometry(16, 30.0, seed=3)
    res = tune(geom, n=16384, steps=32, cells=(1.5, 2.0), margins=(0.8,), ks=(32, 64), device=DEV)
    assert len(res["table"]) == 4
    assert (res["ns_per_spin_step"], res["cell"], res["margin"], res["k"]) == min(res["table"])
    assert res["ns_per_spin_step"] > 0
