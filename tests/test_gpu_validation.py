"""The reference's validation suite on the HIP walker, at the reference's own
size (1e5 spins x 5e4 steps of 2 us in a 10 um box, D = 3 um^2/ms), plus
restricted diffusion in an impermeable cylinder against the Burcaw et al.
(2015) series for the radial D(t).

Ports (same geometries, same assertions):
  tests/validation/narrow_pulse/test_01_free_diffision_empty_arena.py:13-51
  tests/validation/narrow_pulse/test_02_intra_extra_axonal_compartmentation.py:17,81,155,244
  tests/validation/narrow_pulse/test_03_periodic_boundary_xyz.py:15,109
The reference loops sim.step(dt) 5e4 times; ``run(nt, dt)`` advances the same
steps in fewer launches (the per-step API is the same walk:
test_gpu_parity.py::test_free_diffusion_empty_arena_step_api).
Cylinder: the sphere-chain cylinder of the calibration study
(tests/calibration/num_molecules_and_time_step_calibration/
calibration_for_num_molecules_and_time_step_search.py:102-140) and a denser
chain, against calculateDanalytical_longtime (:35-61).
"""
import numpy as np
import pytest
import torch

from nozomi_amd.simulation_engine.diffsim3d import DiffSim3d
from nozomi_amd.simulation_engine.geometry import SimGeometry3D, Structure3D

from .closed_forms import burcaw_dperp, cylinder_chain

pytestmark = pytest.mark.gpu

L, D, T2, RHO = 10.0, 3.0, 100.0, 1
DT, NT, SPINS = 0.002, 50000, 100000


def _inside_any(geom, pts, chunk=4096):
    """Host FP64 union test over every structure but the box."""
    out = np.zeros(pts.shape[1], bool)
    for s in geom.structures[:-1]:
        for i in range(0, pts.shape[1], chunk):
            out[i:i + chunk] |= s.isinside(pts[:, i:i + chunk])
    return out


def _msd_over_2t(sim):
    d = (sim.spins_d - sim.spins0_d).get().astype(np.float64)
    return (d ** 2).sum(axis=1) / sim.nspins / (2 * DT * NT)


def _run(geom, structures, seed):
    sim = DiffSim3d(geom, SPINS)
    sim.set_segments()
    sim.setup(structures=structures, initstates_int=seed)
    print(f"  {SPINS} spins x {NT} steps, seeded in {structures}", flush=True)
    sim.run(NT, DT)
    return sim


def _chain_geometry():
    """20 unit spheres along z with a sinusoidal y offset and 2 + 2 images
    (test_02...py:161-169 and :250-258)."""
    g = SimGeometry3D(L, L, L, D, T2, RHO)
    nsphere = 20
    sz = np.arange(-nsphere / 2 - 2, nsphere / 2 + 3) * L / nsphere
    sy = np.sin(2 * np.pi * np.arange(-2, nsphere + 3) / nsphere) * 0.25
    g.add_structure(Structure3D(np.zeros_like(sz), sy, sz, np.ones_like(sz), D, T2, RHO))
    return g


def _periodic_pair_geometry():
    """Two sinusoidal chains of 200 unit spheres on the x = +-L/2 faces with
    29 + 29 z images each, one structure (test_03...py:27-41, :121-135)."""
    g = SimGeometry3D(L, L, L, D, T2, RHO)
    sz = np.linspace(-L / 2, L / 2, num=200)
    sz = np.concatenate((sz[-30:-1] - L, sz, sz[1:30] + L))
    sy = np.concatenate((np.sin(2 * np.pi * sz / L), np.sin(2 * np.pi * sz / L)))
    sx = np.concatenate((np.full(sz.shape, L / 2), np.full(sz.shape, -L / 2)))
    sz = np.concatenate((sz, sz))
    g.add_structure(Structure3D(sx, sy, sz, np.ones_like(sz), D, T2, RHO))
    return g


def _single_sphere_geometry():
    g = SimGeometry3D(L, L, L, D, T2, RHO)
    g.add_structure(Structure3D(np.array([0.0]), np.array([0.0]), np.array([0.0]),
                                np.array([1.0]), D, T2, RHO))
    return g


def test_01_free_diffusion_empty_arena():
    g = SimGeometry3D(L, L, L, D, T2, RHO)
    sim = _run(g, None, 101)
    p = sim.spins_d.get()
    assert (np.abs(p) <= L / 2).all()
    msd = _msd_over_2t(sim)
    print(f"  D estimate per axis {msd}", flush=True)
    assert np.allclose(msd, D, rtol=1e-1)
    # sharper: the sampling s.e. of MSD/(2t) is D sqrt(2/N) = 0.013
    assert np.abs(msd - D).max() < 5 * D * np.sqrt(2.0 / SPINS)


def test_02_inside_single_sphere():
    g = _single_sphere_geometry()
    sim = _run(g, [0], 102)
    assert g.structures[0].isinside(sim.spins_d.get()).all()
    # long-time MSD per axis inside a sphere of radius 1 -> 2 r^2 / 5
    d = (sim.spins_d - sim.spins0_d).get().astype(np.float64)
    msd = (d ** 2).mean(axis=1)
    print(f"  MSD per axis {msd} (2 r^2/5 = 0.4)", flush=True)
    assert np.all(np.abs(msd / 0.4 - 1) < 0.02)


def test_02_outside_single_sphere():
    g = _single_sphere_geometry()
    sim = _run(g, [1], 103)
    assert not _inside_any(g, sim.spins_d.get()).any()
    assert np.abs(sim.sig_d.get().mean() - np.exp(-NT * DT / T2)) < 1e-2


def test_02_inside_multiple_spheres():
    g = _chain_geometry()
    sim = _run(g, [0], 104)
    assert _inside_any(g, sim.spins_d.get()).all()
    msd = _msd_over_2t(sim)
    print(f"  D estimate per axis {msd}", flush=True)
    assert D - 0.2 < msd[2] < D + 0.2
    assert msd[1] < 0.2 and msd[0] < 0.2
    assert np.abs(sim.sig_d.get().mean() - np.exp(-NT * DT / T2)) < 1e-2


def test_02_outside_multiple_spheres():
    g = _chain_geometry()
    sim = _run(g, [1], 105)
    assert not _inside_any(g, sim.spins_d.get()).any()
    msd = _msd_over_2t(sim)
    print(f"  D estimate per axis {msd}", flush=True)
    assert D - 0.2 < msd[2] < D + 0.2
    assert msd[1] > D - 0.5 and msd[0] > D - 0.5
    assert np.abs(sim.sig_d.get().mean() - np.exp(-NT * DT / T2)) < 1e-2


def test_03_periodic_inside_two_chains():
    g = _periodic_pair_geometry()
    sim = _run(g, [0], 106)
    p = sim.spins_d.get()
    assert _inside_any(g, p).all()
    assert (np.abs(p) <= L / 2).all()
    msd = _msd_over_2t(sim)
    print(f"  D estimate per axis {msd}", flush=True)
    assert D - 0.5 < msd[2] < D + 0.5
    assert msd[1] < 0.2 and msd[0] < 0.2
    assert np.abs(sim.sig_d.get().mean() - np.exp(-NT * DT / T2)) < 1e-2
    # the chains straddle the x faces: spins do cross them
    assert (np.abs(p[0]) > L / 2 - 1.0).mean() > 0.5


def test_03_periodic_outside_two_chains():
    g = _periodic_pair_geometry()
    sim = _run(g, [1], 107)
    assert not _inside_any(g, sim.spins_d.get()).any()
    msd = _msd_over_2t(sim)
    print(f"  D estimate per axis {msd}", flush=True)
    assert D - 0.2 < msd[2] < D + 0.2
    assert msd[1] > D - 0.5 and msd[0] > D - 0.5
    assert np.abs(sim.sig_d.get().mean() - np.exp(-NT * DT / T2)) < 1e-2


# ---- restricted diffusion in a cylinder ------------------------------------
def _cylinder_dperp(g, N, dt, T, seed):
    sim = DiffSim3d(g, N)
    sim.setup(structures=[0], initstates_int=seed)
    sm = torch.zeros((T, 6), dtype=torch.float64, device=sim.device)
    print(f"  cylinder walk: {N} spins x {T} steps of {dt} ms", flush=True)
    sim.run(T, dt, stepmom=sm)
    sm = sm.cpu().numpy()
    t = dt * np.arange(1, T + 1)
    return t, (sm[:, 0] + sm[:, 1]) / (4 * t * N), sim


@pytest.mark.parametrize("build", ["calibration chain (100 spheres, 0.2 um apart)",
                                   "dense chain (0.05 um apart)"])
def test_restricted_cylinder_dperp_burcaw(build):
    """Radial D(t) inside a cylinder of radius a = 0.5 um (D0 = 2), 1e5 spins,
    20 ms of 0.5 us steps, against the closed-form series.  The calibration
    chain (100 spheres over 20 um, calibration...py:115-130) has a rippled
    wall between radius sqrt(a^2 - (s/2)^2) = 0.490 and a, so its D(t) must
    lie between the two cylinders' series; the dense chain (ripple 0.06%)
    must follow the series for radius a to within 2%."""
    a, D0, dt, T, N = 0.5, 2.0, 0.0005, 40000, 100000
    if build.startswith("calibration"):
        g = cylinder_chain(SimGeometry3D, Structure3D, a, 20.0, 20.0 / 99, 30, D0)
        s = 20.0 / 99
        lo_r = np.sqrt(a ** 2 - (s / 2) ** 2)
    else:
        g = cylinder_chain(SimGeometry3D, Structure3D, a, 20.0, 0.05, 60, D0)
        lo_r = a
    t, dp, sim = _cylinder_dperp(g, N, dt, T, 7)
    assert _inside_any(g, sim.spins_d.get()).all()
    # compare at times well resolved by the step (t >= 40 dt) and past the
    # free regime
    idx = np.searchsorted(t, [0.05, 0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10.0, 19.999])
    tt, dd = t[idx], dp[idx]
    hi = burcaw_dperp(tt, a, D0)
    lo = burcaw_dperp(tt, lo_r, D0)
    # sampling s.e. of the mean of (dx^2 + dy^2)/(4t): at long times the
    # displacement is uniform in the disc, rel. s.e. ~ sqrt(2/N) at most
    se = dd * np.sqrt(2.0 / N)
    print("  t (ms)        ", tt)
    print("  D_perp sim    ", dd)
    print("  Burcaw a      ", hi)
    if lo_r != a:
        print(f"  Burcaw {lo_r:.3f}", lo)
    if lo_r == a:
        rel = np.abs(dd / hi - 1)
        print("  rel. diff", rel, flush=True)
        assert (rel < 0.02 + 4 * se / dd).all()
    else:
        assert (dd > lo - 4 * se).all() and (dd < hi * 1.02 + 4 * se).all()
    # long-time limit a^2 / (4 t) (the sum is exp-small by 10 ms)
    assert abs(dd[-1] / (lo_r ** 2 / (4 * tt[-1])) - 1) < 0.06
