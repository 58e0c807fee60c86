"""Pins of the oracle's walk (orc_walk), seeding (orc_seed) and the restated
segment lists against things other than the oracle's own formulas: known
answers assembled from the separately pinned RNG and ray/sphere primitives,
closed forms of diffusion, and brute force.

Each check is chosen so that a plausible mistake fails it: a wrong jump scale
or draw index (free-walk known answer), a wrap that forgets the spins0 shift or
the phase correction or uses the wrong step's area (wrapped free walk with a
balanced waveform), a wrong T2 rule (free and single-collision weights), the
inside/outside tolerance radii or the scatter rule (hand-traced collisions),
leaking walls (confinement), the wrong reflection law (restricted MSD limit).
"""
import numpy as np
import pytest

from nozomi_amd.substrate import single_sphere_geometry
from oracle.reference_segments import reference_segments

TOL = float(np.float32(1e-4))  # sim3d_kernel.cu:34, `const float tol = 1e-4`


def _arena(oracle_lib, spheres, L, D=2.0, T2=100.0):
    sp = np.asarray(spheres, dtype=np.float64).reshape(5, -1)
    ns = int(sp[4].max()) + 1 if sp.shape[1] else 0
    return oracle_lib.OracleSubstrate(sp, ns, np.full(ns + 1, D), np.full(ns + 1, T2), (L, L, L))


def _far_sphere(L):
    # one tiny sphere in a corner, out of reach of the spins below
    return [[L / 2 - 0.01], [L / 2 - 0.01], [L / 2 - 0.01], [1e-3], [0]]


def test_free_walk_known_answer(oracle_lib):
    """Free spins: displacement = sum_t ds * normals(seed, walk stream, spin,
    step t, draw 0); wrapped into [-L/2, L/2] with pos - pos0 continuous; the
    phase moment with a balanced g and area_t = sum_{s<=t} g_s equals
    sum_t g_t * (unwrapped x_t); T2 weight exp(-T dt / T2)."""
    L, D, dt, T, N, seed = 1.5, 2.0, 0.01, 300, 32, 777
    o = _arena(oracle_lib, _far_sphere(L), L, D=D, T2=80.0)
    x0 = np.zeros((3, N))
    x0[0] = np.linspace(-0.3, 0.3, N)   # far from the corner sphere
    pos, pos0, sig = x0.copy(), x0.copy(), np.ones(N)
    g = np.sin(np.linspace(0, 2 * np.pi, T, endpoint=False)) * 0.3
    g -= g.mean()
    area = np.cumsum(g)
    mom = np.zeros((3, N))
    o.walk(pos, pos0, sig, seed=seed, step0=0, nsteps=T, dt=dt, g=g, garea=area, mom=mom)
    ds = np.sqrt(2 * D * dt)
    disp = np.zeros((3, N))
    mom_ref = np.zeros((3, N))
    for i in range(N):
        steps = np.stack([oracle_lib.normals(seed, 0, i, t, 0, 1)[0] for t in range(T)])  # [T, 3]
        traj = np.cumsum(ds * steps, axis=0)
        disp[:, i] = traj[-1]
        mom_ref[:, i] = (g[:, None] * (x0[:, i][None, :] + traj)).sum(0)
    assert np.allclose(pos - pos0, disp, atol=1e-12, rtol=0)
    assert np.all(np.abs(pos) <= L / 2)
    assert np.allclose(pos, x0 + disp - L * np.round((x0 + disp) / L), atol=1e-12)
    # wraps happened (the box is small), so the phase check covers them
    assert np.abs(np.round((x0 + disp) / L)).sum() > 10
    assert np.allclose(mom, mom_ref, atol=1e-10, rtol=0)
    assert np.allclose(sig, np.exp(-T * dt / 80.0), rtol=1e-12)


def test_step_moments_known_answer(oracle_lib):
    """Per-step moment sums (computeDiffusionCoefficientsAndKurtosis,
    sim3d_kernel.cu:162-238, restated at walker_oracle.cpp's stepmom) of free
    spins in a small periodic box: row t holds sum_i d_i(t)^2 per axis, then
    sum_i d_i(t)^4, with d the unwrapped displacement after step t's wrap,
    built here from the RNG known answer.  A wrong step index, axis order,
    power, or a displacement taken before the spins0 shift fails it."""
    L, D, dt, T, N, seed = 1.5, 2.0, 0.01, 120, 24, 4321
    o = _arena(oracle_lib, _far_sphere(L), L, D=D)
    x0 = np.zeros((3, N))
    x0[1] = np.linspace(-0.3, 0.3, N)
    pos, pos0, sig = x0.copy(), x0.copy(), np.ones(N)
    res = o.walk(pos, pos0, sig, seed=seed, step0=0, nsteps=T, dt=dt, stepmom=True)
    ds = np.sqrt(2 * D * dt)
    traj = np.zeros((T, 3, N))
    for i in range(N):
        steps = np.stack([oracle_lib.normals(seed, 0, i, t, 0, 1)[0] for t in range(T)])
        traj[:, :, i] = np.cumsum(ds * steps, axis=0)
    ref = np.concatenate([(traj ** 2).sum(2), (traj ** 4).sum(2)], axis=1)  # [T, 6]
    assert np.abs(np.round((x0 + traj[-1]) / L)).sum() > 5  # wraps happened
    assert np.allclose(res["stepmom"], ref, rtol=1e-12, atol=1e-14)


def test_step_moments_free_diffusion_closed_form(oracle_lib):
    """Empty arena, N spins: E[sum d^2] = N 2 D t and E[sum d^4] = 3 N (2 D t)^2
    per axis at t = (k+1) dt (Gaussian displacements), within 5 standard
    errors (var d^2 = 2 s^4, var d^4 = 96 s^8); the excess kurtosis of the
    reference's host formula K = N S4 / S2^2 - 3 is then near 0."""
    L, D, dt, T, N = 40.0, 2.0, 0.01, 200, 20000
    o = _arena(oracle_lib, _far_sphere(L), L, D=D)
    rng = np.random.default_rng(3)
    x0 = (rng.random((3, N)) - 0.5) * 4.0
    pos, pos0, sig = x0.copy(), x0.copy(), np.ones(N)
    res = o.walk(pos, pos0, sig, seed=99, step0=0, nsteps=T, dt=dt, stepmom=True)
    sm = res["stepmom"]
    s2 = 2 * D * dt * np.arange(1, T + 1)  # variance per axis after each step
    for a in range(3):
        z2 = (sm[:, a] - N * s2) / (np.sqrt(2 * N) * s2)
        z4 = (sm[:, 3 + a] - 3 * N * s2 ** 2) / (np.sqrt(96 * N) * s2 ** 2)
        assert np.abs(z2).max() < 5, np.abs(z2).max()
        assert np.abs(z4).max() < 5, np.abs(z4).max()
    K = N * sm[-1, 3:] / sm[-1, :3] ** 2 - 3
    assert np.abs(K).max() < 0.15, K


def _first_jump(oracle_lib, seed, spin, step, draw, ds):
    return ds * oracle_lib.normals(seed, 0, spin, step, draw, 1)[0]


def test_single_collision_outside_hand_traced(oracle_lib):
    """A box spin aimed at a sphere so that it touches the r + tol shell half
    way: it stops there (f_out = 1/2), scatters with the remaining fraction
    sqrt(1 - 1/4) along the draw-1 direction, weight exp(-dt/T2 * (1/4 + f2))."""
    L, D, dt, T2, r, seed = 10.0, 2.0, 0.02, 50.0, 0.3, 31
    ds = np.sqrt(2 * D * dt)
    c = np.zeros(3)
    checked = 0
    for spin in range(40):
        j1 = _first_jump(oracle_lib, seed, spin, 0, 0, ds)
        s = r + TOL + 0.5 * np.linalg.norm(j1)
        p = c - s * j1 / np.linalg.norm(j1)
        p1 = p + 0.5 * j1
        j2 = _first_jump(oracle_lib, seed, spin, 0, 1, ds)
        f2 = np.sqrt(1 - 0.25)
        # the second leg leaves the shell: |p1 - c + t j2| grows for t > 0
        if (p1 - c) @ j2 < 1e-3 * np.linalg.norm(j2):
            continue
        spheres = [[0.0], [0.0], [0.0], [r], [0]]
        o = _arena(oracle_lib, spheres, L, D=D, T2=T2)
        # the spin index of the walk is its column: place it at column `spin`
        N = spin + 1
        pos = np.zeros((3, N))
        pos[0, :] = 4.0                    # other spins far away
        pos[:, spin] = p
        pos0, sig, nsub = pos.copy(), np.ones(N), np.zeros(N, np.int64)
        o.walk(pos, pos0, sig, seed=seed, step0=0, nsteps=1, dt=dt, nsub=nsub)
        assert nsub[spin] == 1
        assert np.allclose(pos[:, spin], p1 + f2 * j2, atol=1e-12)
        assert sig[spin] == pytest.approx(np.exp(-0.25 * dt / T2 - f2 * dt / T2), rel=1e-12)
        checked += 1
        if checked == 3:
            break
    assert checked == 3


def test_single_collision_inside_hand_traced(oracle_lib):
    """A spin at the centre of a sphere smaller than its jump exits to the
    r - tol shell (f_in = (r - tol) / |j|) and scatters inward."""
    L, D, dt, seed = 10.0, 2.0, 0.02, 32
    ds = np.sqrt(2 * D * dt)
    checked = 0
    for spin in range(60):
        j1 = _first_jump(oracle_lib, seed, spin, 0, 0, ds)
        n1 = np.linalg.norm(j1)
        r = 0.6 * n1
        fin = (r - TOL) / n1
        p1 = fin * j1
        j2 = _first_jump(oracle_lib, seed, spin, 0, 1, ds)
        f2 = np.sqrt(1 - fin * fin)
        e = p1 + f2 * j2
        if np.linalg.norm(e) > r - TOL - 1e-3:  # the second leg must stay inside
            continue
        o = _arena(oracle_lib, [[0.0], [0.0], [0.0], [r], [0]], L, D=D)
        N = spin + 1
        pos = np.zeros((3, N))
        pos[0, :] = 4.0
        pos[:, spin] = 0.0
        pos0, sig, nsub = pos.copy(), np.ones(N), np.zeros(N, np.int64)
        o.walk(pos, pos0, sig, seed=seed, step0=0, nsteps=1, dt=dt, nsub=nsub)
        assert nsub[spin] == 1
        assert np.allclose(pos[:, spin], e, atol=1e-12)
        checked += 1
        if checked == 3:
            break
    assert checked == 3


def test_confinement_and_restricted_msd(oracle_lib):
    """Spins seeded in an impermeable sphere stay within r + tol; long-time
    MSD per axis tends to 2 r^2 / 5 (two independent uniform points in a
    ball); spins seeded outside never get deeper than r - tol."""
    g = single_sphere_geometry(L=6.0, r=1.0, D=2.0)
    o = oracle_lib.OracleSubstrate(g.spheres, g.nstructures, g.structure_D(), g.structure_T2(),
                                   (g.Lx, g.Ly, g.Lz))
    N = 6000
    pin = o.seed(N, [0], seed=3)
    pos, pos0, sig = pin.copy(), pin.copy(), np.ones(N)
    for _ in range(4):
        o.walk(pos, pos0, sig, seed=9, step0=_ * 500, nsteps=500, dt=0.01)
        assert np.linalg.norm(pos, axis=0).max() <= 1.0 + TOL + 1e-9
    msd = ((pos - pos0) ** 2).mean(1)
    assert np.allclose(msd, 0.4, rtol=0.05), msd
    pout = o.seed(3000, [1], seed=4)
    pos, pos0 = pout.copy(), pout.copy()
    o.walk(pos, pos0, np.ones(3000), seed=10, step0=0, nsteps=1000, dt=0.01)
    d = np.linalg.norm(pos - 6.0 * np.round(pos / 6.0), axis=0)
    assert d.min() >= 1.0 - TOL - 1e-9


def test_seed_acceptance_and_uniformity(oracle_lib):
    """Every seeded point lies in a target structure and in no lower-id
    non-target one (geometry.py:127-174); box seeds are uniform."""
    L = 8.0
    spheres = [[0.0, 0.8, 2.0], [0.0, 0.0, 2.0], [0.0, 0.0, 0.0], [1.0, 1.0, 0.7], [0, 1, 2]]
    o = _arena(oracle_lib, spheres, L)
    sp = np.asarray(spheres)
    for targets in ([1], [0, 2], [3]):
        p = o.seed(20000, targets, seed=11)
        inside = np.stack([np.linalg.norm(p - sp[:3, k:k + 1], axis=0) < sp[3, k] for k in range(3)])
        sid = np.where(inside.any(0), np.argmax(inside, axis=0), 3)
        assert set(np.unique(sid)) <= set(targets)
        assert np.all(np.abs(p) < L / 2)
    p = o.seed(200000, [3], seed=12)
    for a in range(3):
        # outside the spheres, still roughly uniform over the box
        assert abs(p[a].mean()) < 0.1
        assert p[a].var() == pytest.approx(L * L / 12, rel=0.05)


def test_candidate_modes_agree(oracle_lib):
    """Exact bins (mode 0) and the reference's padded segment lists (mode 1)
    give bit-identical walks when the lists hold every sphere a jump reaches."""
    from nozomi_amd.substrate import generate_beaded_axons, load_geometry
    M, L = generate_beaded_axons(6, 16.0, seed=2)
    g = load_geometry(M, L, D=2.0, T2=100.0)
    o = oracle_lib.OracleSubstrate(g.spheres, g.nstructures, g.structure_D(), g.structure_T2(),
                                   (L, L, L))
    N = 3000
    p = o.seed(N, np.arange(g.nstructures + 1), seed=6)
    a, a0 = p.copy(), p.copy()
    o.walk(a, a0, np.ones(N), seed=8, step0=0, nsteps=200, dt=0.01)
    seg, sps = reference_segments(g.spheres, (L, L, L), (8, 8, 8))
    o.set_reference_segments((8, 8, 8), seg)
    b, b0 = p.copy(), p.copy()
    o.walk(b, b0, np.ones(N), seed=8, step0=0, nsteps=200, dt=0.01)
    same = np.all(a == b, axis=0) & np.all(a0 == b0, axis=0)
    assert same.mean() > 0.99


def test_reference_segments_brute_force():
    """The restated segment lists hold exactly the spheres whose AABB comes
    within 1 um of the segment (diffsim3d.py:128-141), padded to the max."""
    rng = np.random.default_rng(1)
    S, L, n = 60, 10.0, (3, 4, 2)
    sp = np.zeros((5, S))
    sp[:3] = rng.uniform(-L / 2, L / 2, (3, S))
    sp[3] = rng.uniform(0.2, 1.0, S)
    seg, sps = reference_segments(sp, (L, L, L), n)
    d = [L / k for k in n]
    counts = []
    for p_ in range(n[2]):
        for m in range(n[1]):
            for q in range(n[0]):
                lo = np.array([q * d[0], m * d[1], p_ * d[2]]) - L / 2
                hi = lo + np.array(d)
                # signed distance from the sphere's AABB to the segment box
                gap = np.maximum(lo[:, None] - (sp[:3] + sp[3]), (sp[:3] - sp[3]) - hi[:, None]).max(0)
                want = set(np.flatnonzero(gap < 1.0))
                got = seg[sps * (p_ * n[0] * n[1] + m * n[0] + q):][:sps]
                assert want <= set(got.tolist())
                assert len(set(got.tolist())) == sps
                counts.append(len(want))
    assert sps == max(counts)


def test_restricted_cylinder_matches_burcaw(oracle_lib):
    """Spins inside a cylinder of radius a = 0.5 um (a dense sphere chain,
    wall ripple 0.06%): the oracle's radial D(t) follows the Burcaw et al.
    series (closed form, tests/closed_forms.py) and reaches its long-time
    limit a^2/(4t).  4000 spins, 10 ms of 1 us steps; sampling s.e. about
    2%, so 8% bars."""
    from nozomi_amd.simulation_engine.geometry import SimGeometry3D, Structure3D
    from .closed_forms import burcaw_dperp, cylinder_chain
    a, D0, dt, T, N = 0.5, 2.0, 0.001, 10000, 4000
    g = cylinder_chain(SimGeometry3D, Structure3D, a, 6.0, 0.05, 60, D0)
    o = oracle_lib.OracleSubstrate(g.spheres, g.nstructures, g.structure_D(), g.structure_T2(),
                                   (g.Lx, g.Ly, g.Lz))
    p0 = o.seed(N, [0], seed=12)
    pos, pos0, sig = p0.copy(), p0.copy(), np.ones(N)
    sm = o.walk(pos, pos0, sig, seed=13, step0=0, nsteps=T, dt=dt, stepmom=True)["stepmom"]
    assert g.structures[0].isinside(pos).all()
    t = dt * np.arange(1, T + 1)
    dp = (sm[:, 0] + sm[:, 1]) / (4 * t * N)
    idx = np.searchsorted(t, [0.05, 0.2, 1.0, 5.0, 9.999])
    ref = burcaw_dperp(t[idx], a, D0)
    rel = np.abs(dp[idx] / ref - 1)
    assert (rel < 0.08).all(), (t[idx], dp[idx], ref)
    assert abs(dp[-1] / (a ** 2 / (4 * t[-1])) - 1) < 0.08
    # axial diffusion is free
    assert abs(sm[-1, 2] / (2 * t[-1] * N) / D0 - 1) < 0.1
