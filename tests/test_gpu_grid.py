"""The cell grid built on the device (nz_substrate_create) equals the host
builder's (nz_grid_build.h, run here through the CPU build of the walk core)
entry for entry: CSR starts, sort ranks and the 32-byte records bit for bit,
plus the summary figures of nz_substrate_info."""
import numpy as np
import pytest
import torch

from nozomi_amd.substrate import (cylinder_geometry, generate_beaded_axons, load_geometry,
                                  single_sphere_geometry)
from tests.cpu_build import CpuWalker

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _beaded(n, L, seed=20261203):
    M, Lb = generate_beaded_axons(n, L, seed=seed)
    return load_geometry(M, Lb, D=2.0, T2=100.0)


CASES = {
    "16 axons, 30 um, 1.25 / 0.8": (lambda: _beaded(16, 30.0), 1.25, 0.8),
    "16 axons, 30 um, 2.0 / 1.0": (lambda: _beaded(16, 30.0), 2.0, 1.0),
    "16 axons, 30 um, 1.1 / 0.7 (partial regions)": (lambda: _beaded(16, 30.0), 1.1, 0.7),
    "200 axons, 60 um, 1.25 / 0.8": (lambda: _beaded(200, 60.0), 1.25, 0.8),
    "single sphere": (lambda: single_sphere_geometry(), 1.0, 0.8),
    "cylinder": (lambda: cylinder_geometry(), 1.25, 0.8),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_grid_equals_host_grid(name):
    make, cell, margin = CASES[name]
    geom = make()
    print(f"case {name}: {geom.spheres.shape[1]} spheres", flush=True)
    sub = geom.device_substrate(cell_size=cell, margin=margin, device=DEV)
    start, rank, cand = sub.grid()
    inf = sub.info()
    hs, hr, hc = CpuWalker(geom, cell=cell, margin=margin).grid()
    assert np.array_equal(start, hs)
    assert np.array_equal(rank, hr)
    assert cand.shape == hc.shape
    assert np.array_equal(cand.view(np.uint32), hc.view(np.uint32))
    # reach entries of the first pass
    rs, reach = sub.reach()
    hrs, hreach = CpuWalker(geom, cell=cell, margin=margin).reach()
    assert np.array_equal(rs, hrs)
    assert np.array_equal(reach, hreach)
    # summary figures
    lens = np.diff(hs)
    assert inf["ncand_total"] == hs[-1]
    assert inf["ncand_max"] == lens.max()
    n = inf["ncell"]
    box = lens.reshape(n[2] + 2, n[1] + 2, n[0] + 2)[1:-1, 1:-1, 1:-1]
    assert inf["ncand_mean_box"] == pytest.approx(box.mean(), rel=1e-12)
    # ranks are a permutation of the in-box cells
    assert np.array_equal(np.sort(rank), np.arange(rank.size))


def test_dense_cell_lists_walk_in_fp64(oracle_lib):
    """300 concentric spheres of one structure: a cell lists all of them,
    beyond the 255 entries the FP32 scan keys; such a cell is scanned in FP64
    and the walk still follows the oracle."""
    from nozomi_amd.simulation_engine.diffsim3d import DiffSim3d
    from nozomi_amd.simulation_engine.geometry import SimGeometry3D, Structure3D
    g = SimGeometry3D(10.0, 10.0, 10.0, 2.0, 100.0, 1)
    r = np.linspace(0.5, 1.5, 300)
    z = np.zeros_like(r)
    g.add_structure(Structure3D(z, z, z, r, 2.0, 100.0, 1))
    N, T, dt = 2048, 100, 0.02
    sim = DiffSim3d(g, N, cell_size=1.0, margin=0.8, counters=True)
    sim.setup(initstates_int=21)
    assert sim.sub.info()["ncand_max"] > 255
    p0 = sim.spins_d.get().astype(np.float64)
    sim.run(T, dt)
    gp, gd = sim.walker.read_f64()
    _, nguard = sim.walker.counters()
    orc = oracle_lib.OracleSubstrate(g.spheres, g.nstructures, g.structure_D(), g.structure_T2(),
                                     (10.0, 10.0, 10.0))
    pos, pos0 = p0.copy(), p0.copy()
    orc.walk(pos, pos0, np.ones(N), seed=21, step0=0, nsteps=T, dt=dt)
    d = gp - pos
    err = np.abs(d - 10.0 * np.round(d / 10.0)).max(axis=0)
    print(f"dense cells: FP64 scans {int(nguard.sum())}, within 1e-5 {np.mean(err < 1e-5):.4f}, "
          f"bit-identical {np.mean(err == 0):.4f}", flush=True)
    assert nguard.sum() > 1000
    assert np.mean(err < 1e-5) >= 0.999
    # the union of the spheres confines the spins seeded inside it
    inside0 = np.sqrt((p0 ** 2).sum(0)) < 1.5
    assert (np.sqrt((gp[:, inside0] ** 2).sum(0)) < 1.5).all()
