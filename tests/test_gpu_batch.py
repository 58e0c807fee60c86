"""Many simulations per launch (nz_batch_*) against the single walker: every
simulation's trajectories, T2 weights and phase moments are bit-identical to
the walker's run alone on the same substrate, key and start positions, with
or without the captured graph, while other simulations share the launches
and come and go."""
import numpy as np
import pytest
import torch

from nozomi_amd import runtime as rt
from nozomi_amd.batch import Batch, BatchFull
from nozomi_amd.protocol import PGSEShells, icosahedron_6
from nozomi_amd.substrate import generate_beaded_axons, load_geometry, single_sphere_geometry

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def subs():
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    g0 = load_geometry(M, L, D=2.0, T2=100.0)
    g1 = single_sphere_geometry(L=10.0, r=1.0, D=3.0, T2=50.0)
    return (g0, g0.device_substrate(device=DEV)), (g1, g1.device_substrate(device=DEV))


def _spec(subs):
    (g0, s0), (g1, s1) = subs
    prot = PGSEShells(2.0, 4.0, 0.02, [1.0], icosahedron_6())
    g = torch.from_numpy(prot.g).to(DEV)
    ga = torch.from_numpy(prot.garea).to(DEV)
    # (substrate, geometry, n, walk key, steps, dt, phase, T2, start-seed)
    return [
        (s0, g0, 5000, 11, prot.nsteps, prot.dt, (g, ga), False, 1),
        (s0, g0, 3000, 12, 200, 0.02, None, True, 2),
        (s1, g1, 2000, 13, 250, 0.01, None, True, 3),
    ]


def _alone(sub, n, key, T, dt, ph, t2, start):
    w = rt.Walker(sub, n)
    w.set_positions(start)
    g, ga = ph if ph else (None, None)
    w.walk(key, T, dt, g=g, garea=ga, t2=t2)
    return _read_walker(w, n)


def _read_walker(w, n):
    out = [torch.empty((3, n), device=DEV), torch.empty((3, n), device=DEV),
           torch.empty(n, device=DEV), torch.empty((3, n), device=DEV)]
    w.read(*out)
    return [o.cpu() for o in out]


def _read_batch(b, sim, n):
    out = [torch.empty((3, n), device=DEV), torch.empty((3, n), device=DEV),
           torch.empty(n, device=DEV), torch.empty((3, n), device=DEV)]
    b.read(sim, n, *out)
    return [o.cpu() for o in out]


def _starts(spec):
    return [sub.seed(n, targets=np.arange(geo.nstructures + 1), seed=s) for sub, geo, n, *_, s in spec]


def test_batch_matches_walker_alone(subs):
    spec = _spec(subs)
    starts = _starts(spec)
    b = Batch(16384, max_sims=8, device=DEV)
    sims = []
    for (sub, geo, n, key, T, dt, ph, t2, _), p in zip(spec, starts):
        g, ga = ph if ph else (None, None)
        sims.append(b.add(sub, n, key, T, dt, g=g, garea=ga, t2=t2, k=64, pos=p))
    for it in range(20):
        print(f"batch launch pair {it}", flush=True)
        b.run(2)
        if all(b.status(s)[1] for s in sims):
            break
    for (sub, geo, n, key, T, dt, ph, t2, _), p, sim in zip(spec, starts, sims):
        assert b.status(sim) == (T, True)
        ref = _alone(sub, n, key, T, dt, ph, t2, p)
        got = _read_batch(b, sim, n)
        for name, r, x in zip(("pos", "pos0", "sig", "mom"), ref, got):
            assert torch.equal(r, x), f"sim {sim} {name} differs"
    b.close()


def test_batch_graph_replays_match_eager(subs):
    """The captured launch pair is replayed for changing members (eager adds,
    removes and synchronisation between replays); each replay's finished
    simulations equal those of a batch that launches eagerly.  A snapshot in
    mid-walk is not compared: spins may run ahead of a launch's target by an
    amount that depends on their wavefront's composition, which the counting
    sort does not fix."""
    (g0, s0), (g1, s1) = subs
    bg = Batch(8192, max_sims=4, device=DEV, graph=True)
    be = Batch(8192, max_sims=4, device=DEV, graph=False)
    prot = PGSEShells(0.4, 0.8, 0.02, [1.0], icosahedron_6())
    g = torch.from_numpy(prot.g).to(DEV)
    ga = torch.from_numpy(prot.garea).to(DEV)
    k = (prot.nsteps + 1) // 2  # one replay (two launches) completes a simulation
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        jobs = []
        for q, (sub, geo, n) in enumerate(((s0, g0, 3000), (s1, g1, 1000))):
            p = sub.seed(n, targets=np.arange(geo.nstructures + 1), seed=100 + 10 * rep + q)
            key = 1000 + 10 * rep + q
            jobs.append((bg.add(sub, n, key, prot.nsteps, prot.dt, g=g, garea=ga, k=k, pos=p),
                         be.add(sub, n, key, prot.nsteps, prot.dt, g=g, garea=ga, k=k, pos=p), n))
        bg.run(2)
        be.run(2)
        torch.cuda.synchronize()
        for sg, se, n in jobs:
            assert bg.status(sg) == (prot.nsteps, True) and be.status(se) == (prot.nsteps, True)
            a, e = _read_batch(bg, sg, n), _read_batch(be, se, n)
            for x, y in zip(a, e):
                assert torch.equal(x, y), f"replay {rep} differs"
            bg.remove(sg)
            be.remove(se)
    bg.close()
    be.close()


def test_batch_admission_reuse_and_signals(subs):
    (g0, s0), _ = subs
    b = Batch(3 * 256, max_sims=4, device=DEV)
    a = b.add(s0, 500, 21, 100, 0.02, pos=s0.seed(500, targets=[g0.nstructures], seed=7))  # 2 blocks
    with pytest.raises(BatchFull):
        b.add(s0, 300, 22, 100, 0.02, targets=[g0.nstructures])        # needs 2, 1 free
    c = b.add(s0, 200, 23, 100, 0.02, targets=[g0.nstructures])        # fits the last block
    assert b.free()[1] == 0
    b.remove(a)
    p = s0.seed(400, targets=np.arange(g0.nstructures), seed=8)
    prot = PGSEShells(1.0, 1.5, 0.02, [1.0, 2.0], icosahedron_6())
    g = torch.from_numpy(prot.g).to(DEV)
    ga = torch.from_numpy(prot.garea).to(DEV)
    d = b.add(s0, 400, 24, prot.nsteps, prot.dt, g=g, garea=ga, k=32, pos=p)  # reuses a's blocks
    while not (b.status(d)[1] and b.status(c)[1]):
        b.run(2)
    ref = _alone(s0, 400, 24, prot.nsteps, prot.dt, (g, ga), False, p)
    got = _read_batch(b, d, 400)
    for r, x in zip(ref, got):
        assert torch.equal(r, x)
    dirs = torch.from_numpy(prot.dirs.astype(np.float32)).to(DEV)
    sc = torch.from_numpy(prot.scales.astype(np.float32)).to(DEV)
    sb = b.signals(d, dirs, sc).cpu().numpy()
    sw = rt.signal_sums(ref[3].to(DEV), dirs, sc).cpu().numpy()
    # same moments, summed in another order (FP32 partial sums of 32, then FP64)
    assert np.abs(sb - sw).max() / 400 < 1e-6
    b.close()
