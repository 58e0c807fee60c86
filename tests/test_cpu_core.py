"""The walk kernel's per-spin logic (CPU build of nz_walk_core.h) against the
FP64 oracle, on the small-substrate configuration (16 beaded axons in a 30 um
periodic cube, dt = 20 us), same start positions and Philox stream.

Trajectory agreement is limited by the dynamics, not by arithmetic: at every
wall collision a perturbation of the contact point is amplified (the end of a
chord depends on its start through 1/cos of the incidence angle), so any two
implementations that round differently separate exponentially in the number
of collisions.  test_oracle_is_chaotic shows this on the FP64 oracle itself;
the bar for the kernel is therefore set per collision count.
"""
import numpy as np
import pytest

from nozomi_amd.substrate import generate_beaded_axons, load_geometry, single_sphere_geometry
from nozomi_amd.simulation_engine.geometry import SimGeometry3D
from tests.cpu_build import CpuWalker


def _min_image(d, L):
    return d - L * np.round(d / L)


@pytest.fixture(scope="module")
def c0(oracle_lib):
    M, L = generate_beaded_axons(16, 30.0, seed=20261203)
    g = load_geometry(M, L, D=2.0, T2=100.0)
    o = oracle_lib.OracleSubstrate(g.spheres, g.nstructures, g.structure_D(), g.structure_T2(), (L, L, L))
    return g, o


def _run_pair(g, o, N, nsteps, dt=0.02, seed=1234, targets=None):
    L = g.Lx
    p = o.seed(N, targets if targets is not None else np.arange(g.nstructures + 1), seed=5)
    p = p.astype(np.float32)
    cw = CpuWalker(g)
    cw.init(p)
    out = cw.walk(seed, 0, nsteps, dt)
    pos = p.astype(np.float64)
    pos0 = pos.copy()
    sig = np.ones(N)
    nsub = np.zeros(N, np.int64)
    o.walk(pos, pos0, sig, seed=seed, step0=0, nsteps=nsteps, dt=dt, nsub=nsub)
    err = np.maximum(np.abs(_min_image(out["pos_f64"] - pos, L)).max(0),
                     np.abs(out["disp_f64"] - (pos - pos0)).max(0))
    return out, err, nsub, sig


def test_free_diffusion_bitlevel(oracle_lib):
    g = SimGeometry3D(30.0, 30.0, 30.0, 2.0, 100.0, 1)
    o = oracle_lib.OracleSubstrate(g.spheres, 0, g.structure_D(), g.structure_T2(), (30.0, 30.0, 30.0))
    out, err, nsub, _ = _run_pair(g, o, 2048, 200)
    assert err.max() < 1e-12
    assert (out["nsub"] == 0).all()


def test_single_sphere_short(oracle_lib):
    g = single_sphere_geometry(L=10.0, r=1.0, D=2.0)
    o = oracle_lib.OracleSubstrate(g.spheres, 1, g.structure_D(), g.structure_T2(), (10.0, 10.0, 10.0))
    for targets in ([0], [1]):
        out, err, nsub, _ = _run_pair(g, o, 2048, 10, targets=targets)
        assert (out["nsub"] == nsub).all()
        assert err.max() < 1e-8, err.max()


def test_trajectories_1000_steps(c0):
    g, o = c0
    out, err, nsub, sig = _run_pair(g, o, 4096, 1000)
    same = out["nsub"] == nsub
    few = nsub <= 60
    print(f"within 1e-5: all {np.mean(err < 1e-5):.4f}; <=60 collisions {np.mean(err[few] < 1e-5):.4f}; "
          f"branch-history mismatches {np.sum(~same)}; FP64 re-solves {out['nguard'].sum()}")
    assert np.mean(err[few] < 1e-5) > 0.995
    assert np.mean(err < 1e-5) > 0.85
    assert np.median(err[few]) < 1e-9
    s = np.exp(out["logsig"].astype(np.float64))
    assert np.abs(s[few] / sig[few] - 1).max() < 1e-5


def test_oracle_is_chaotic(c0):
    """The FP64 oracle against itself with start positions perturbed by
    1e-9 um: spins with many wall collisions separate beyond 1e-5 um."""
    g, o = c0
    N = 2048
    p = o.seed(N, np.arange(g.nstructures + 1), seed=5)
    q = p + np.random.default_rng(0).normal(size=p.shape) * 1e-9
    res = []
    for start in (p, q):
        pos = start.copy()
        pos0 = start.copy()
        nsub = np.zeros(N, np.int64)
        o.walk(pos, pos0, np.ones(N), seed=1234, step0=0, nsteps=1000, dt=0.02, nsub=nsub)
        res.append((pos, nsub))
    err = np.abs(_min_image(res[0][0] - res[1][0], g.Lx)).max(0)
    many = res[0][1] > 200
    assert many.sum() > 20
    assert np.mean(err[many] > 1e-5) > 0.5        # collision-rich spins diverge
    assert np.mean(err[res[0][1] <= 10] < 1e-7) > 0.99   # collision-poor spins do not


def test_fp32_state_breaks_the_trajectory_bar(c0):
    """What an FP32 walk state would cost: the FP64 oracle with its positions
    rounded to binary32 after every step (a milder error than FP32 arithmetic
    throughout) against the plain FP64 oracle over 1000 steps.  Most
    collision-rich spins end more than 1e-5 um apart, the per-spin bar the
    FP64-state kernel meets bit for bit."""
    g, o = c0
    N, T, dt, seed = 2048, 1000, 0.02, 1234
    p = o.seed(N, np.arange(g.nstructures + 1), seed=5).astype(np.float32).astype(np.float64)
    ref, ref0 = p.copy(), p.copy()
    nsub = np.zeros(N, np.int64)
    o.walk(ref, ref0, np.ones(N), seed=seed, step0=0, nsteps=T, dt=dt, nsub=nsub)
    pos, pos0 = p.copy(), p.copy()
    for t in range(T):
        o.walk(pos, pos0, np.ones(N), seed=seed, step0=t, nsteps=1, dt=dt)
        pos = pos.astype(np.float32).astype(np.float64)
    err = np.abs(_min_image(pos - ref, g.Lx)).max(0)
    many = nsub > 200
    frac_all, frac_many = np.mean(err < 1e-5), np.mean(err[many] < 1e-5)
    print(f"FP32 state: within 1e-5 um after {T} steps: all {frac_all:.3f}, "
          f"spins with > 200 sub-steps {frac_many:.3f} ({many.sum()} spins)")
    assert many.sum() > 20
    assert frac_many < 0.5
    assert frac_all < 0.9


def _probes(g, rng, n_u, n_s):
    """Uniform positions and positions on and near sphere surfaces (inside
    and outside the tolerance shells), wrapped into the box."""
    L = g.Lx
    sp = g.spheres
    pos_u = rng.uniform(-L / 2, L / 2, size=(3, n_u))
    k = rng.integers(0, sp.shape[1], n_s)
    d = rng.normal(size=(3, n_s))
    d /= np.linalg.norm(d, axis=0)
    shell = sp[3, k] + rng.choice([-1.0, 1.0], n_s) * 10.0 ** rng.uniform(-7, -3, n_s)
    pos = np.concatenate([pos_u, sp[:3, k] + d * shell], axis=1)
    pos = pos - L * np.round(pos / L)
    return np.clip(pos, -L / 2 + 1e-6, L / 2 - 1e-6)


@pytest.mark.parametrize("cell", [1.25, 4.0])
def test_reach_entries_round_as_numpy_float16(c0, cell):
    """Every reach entry of the host grid: h = -2 sigma c rounded to the
    nearest binary16 (numpy's float16 conversion from FP64, ties to even),
    and k the largest binary16 at or below sigma (|c'|^2 - (1 + lambda)
    R'^2) (c' the centre h encodes, R' = r + tol + |c - c'|) less the FP64
    round-off allowance.  Lists start at even entries; the gaps and the
    padding are zero."""
    g, _ = c0
    cw = CpuWalker(g, cell=cell, margin=0.8)
    lam, sigma, lam1, eps = cw.reach_params()
    assert sigma == 2.0 ** np.round(np.log2(sigma)) and lam1 == np.float32(sigma * (1 + 1 / lam))
    start, _, cand = cw.grid()
    rs, reach = cw.reach()
    ncells = rs.size
    lens = np.diff(start)[:ncells]
    assert np.all(rs % 2 == 0) and np.array_equal(np.diff(rs), (lens[:-1] + 1) & ~1)
    cell_of = np.repeat(np.arange(ncells), lens)
    e = np.arange(cell_of.size)
    pos = rs[cell_of] + (e - start[cell_of])
    used = np.zeros(reach.shape[0], bool)
    used[pos] = True
    assert not reach[~used].any()
    nloc = round(ncells ** (1 / 3))
    h = g.Lx / (nloc - 2)
    lx, ly, lz = cell_of % nloc, (cell_of // nloc) % nloc, cell_of // nloc ** 2
    lo = np.stack([-g.Lx / 2 + (l - 1) * h for l in (lx, ly, lz)])
    idx = cand[1::2, 3].view(np.int32)[: cell_of.size]
    sp = g.spheres
    c = sp[:3, idx] - lo
    w = reach[pos]
    hbits = np.stack([w[:, 0] & 0xFFFF, w[:, 0] >> 16, w[:, 1] & 0xFFFF]).astype(np.uint16)
    assert np.array_equal(hbits, (-2.0 * sigma * c).astype(np.float16).view(np.uint16))
    ct = hbits.view(np.float16).astype(np.float64) / (-2.0 * sigma)
    rp = sp[3, idx] + 1e-4 + np.sqrt(((c - ct) ** 2).sum(0)) * (1 + 1e-9)
    ct2, q = (ct ** 2).sum(0), (1 + lam) * rp * rp
    bound = sigma * (ct2 - q) - 1e-12 * sigma * (ct2 + q)
    k = (w[:, 1] >> 16).astype(np.uint16).view(np.float16)
    assert np.all(k.astype(np.float64) <= bound)
    up = np.nextafter(k, np.float16(np.inf)).astype(np.float64)
    assert np.all(up > bound)


@pytest.mark.parametrize("cell", [1.25, 2.0, 4.0])
def test_reach_bound_keeps_every_touchable_sphere(c0, cell):
    """The FP32 reach test (three FMAs on the FP16 reach entry) passes every
    candidate whose R-ball (R = r + tol) the segment touches, judged by the
    FP64 distance from the segment to the centre: segments from 1e-4 um to
    the list margin, probes on and near sphere surfaces.  The pass set stays
    close to the touch set."""
    g, _ = c0
    cw = CpuWalker(g, cell=cell, margin=0.8)
    rng = np.random.default_rng(11)
    pos = _probes(g, rng, 20000, 40000)
    n = pos.shape[1]
    jd = rng.normal(size=(3, n))
    jd /= np.linalg.norm(jd, axis=0)
    jump = jd * 10.0 ** rng.uniform(-4, np.log10(0.7), n)
    jr = 10.0 ** rng.uniform(-4, 0, n)
    out = cw.reach_check(pos, jump, jr)
    print(f"touching {out[:, 1].sum()}, passing {out[:, 2].sum()} "
          f"({out[:, 2].sum() / max(out[:, 1].sum(), 1):.3f}x)")
    assert out[:, 0].sum() == 0
    assert out[:, 2].sum() <= 1.6 * out[:, 1].sum()


@pytest.mark.parametrize("cell", [1.25, 4.0])
def test_two_phase_scan_equals_full_scan(c0, cell):
    """The reach pass may only skip spheres that cannot change the sub-step:
    against visiting every candidate of the list, the exit key and the
    structure id are equal, the entry key is equal whenever either is a root
    the sub-step can take (entries beyond jr belong to skipped obstacles), and
    guards differ only by skipped out-of-reach ones.  Uniform probes and
    probes on and near sphere surfaces (inside and outside the tolerance
    shells), segments from 1e-4 um to the list margin, every remaining
    fraction, every assumed structure id.  4 um cells give lists of up to 66
    entries: every tail length of the eight-candidate reach trips, and lists
    past the 64-entry two-phase limit."""
    g, _ = c0
    cw = CpuWalker(g, cell=cell, margin=0.8)
    rng = np.random.default_rng(7)
    L = g.Lx
    sp = g.spheres
    n_u, n_s = 20000, 60000
    pos_u = rng.uniform(-L / 2, L / 2, size=(3, n_u))
    k = rng.integers(0, sp.shape[1], n_s)
    d = rng.normal(size=(3, n_s))
    d /= np.linalg.norm(d, axis=0)
    shell = sp[3, k] + rng.choice([-1.0, 1.0], n_s) * 10.0 ** rng.uniform(-7, -3, n_s)
    pos_s = sp[:3, k] + d * shell
    pos = np.concatenate([pos_u, pos_s], axis=1)
    pos = pos - L * np.round(pos / L)                     # inside the box
    pos = np.clip(pos, -L / 2 + 1e-6, L / 2 - 1e-6)
    n = pos.shape[1]
    jd = rng.normal(size=(3, n))
    jd /= np.linalg.norm(jd, axis=0)
    jump = jd * 10.0 ** rng.uniform(-4, np.log10(0.7), n)
    jr = 10.0 ** rng.uniform(-4, 0, n)
    sid = rng.integers(0, g.nstructures + 1, n).astype(np.float32)
    out = cw.scan_compare(pos, jump, jr, sid)
    assert out.sum(0).tolist() == [0, 0, 0, 0]
