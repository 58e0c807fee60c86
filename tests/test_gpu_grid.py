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
    # summary figures
    lens = np.diff(hs)
    assert inf["ncand_total"] == hs[-1]
    assert inf["ncand_max"] == lens.max()
    n = inf["ncell"]
    box = lens.reshape(n[2] + 2, n[1] + 2, n[0] + 2)[1:-1, 1:-1, 1:-1]
    assert inf["ncand_mean_box"] == pytest.approx(box.mean(), rel=1e-12)
    # ranks are a permutation of the in-box cells
    assert np.array_equal(np.sort(rank), np.arange(rank.size))


def test_list_overflow_is_reported():
    # 300 concentric spheres: one cell would list all of them (> 255)
    sp = np.zeros((5, 300))
    sp[3] = np.linspace(0.5, 1.5, 300)
    from nozomi_amd.runtime import DeviceSubstrate
    with pytest.raises(ValueError, match="more than 255"):
        DeviceSubstrate(sp, 1, [2.0, 2.0], [100.0, 100.0], (10.0, 10.0, 10.0), cell_size=1.0,
                        margin=0.8, device=DEV)
