"""Substrate feed (the reference's pickle -> fibers -> z padding -> structures
pipeline, utils/common_utils.py:41-55, utils/adjust_geometry.py:3-36) and the
seeded beaded-axon generator."""
import numpy as np
import pytest

from nozomi_amd import substrate as sb
from nozomi_amd.simulation_engine.geometry import SimGeometry3D, Structure3D


NZ = 41  # spheres per fiber (the z padding needs at least 25)


def _two_fibers(L=10.0):
    z = np.linspace(-L / 2, L / 2, NZ)
    f0 = np.stack([np.full(NZ, 1.0), np.full(NZ, 2.0), z, np.full(NZ, 0.7), np.zeros(NZ)], 1)
    f1 = np.stack([np.full(NZ, -3.0), np.full(NZ, -1.0), z, np.full(NZ, 0.5), np.ones(NZ)], 1)
    return np.concatenate([f0, f1]), L


def test_split_at_box_top():
    M, L = _two_fibers()
    fibers = sb.split_matrix_to_list(M, L)
    assert len(fibers) == 2
    assert all(f[-1, 2] == L / 2 for f in fibers)
    assert sum(len(f) for f in fibers) == len(M)


def test_create_boundary_z_pads_25_per_end():
    M, L = _two_fibers()
    fibers = sb.createBoundaryZ(sb.split_matrix_to_list(M, L), L)
    for f in fibers:
        assert len(f) == NZ + 50
        # prepended images are the fiber's last 25 spheres shifted down by L
        assert np.allclose(f[:25, 2], f[25 + NZ - 25:25 + NZ, 2] - L)
        assert np.allclose(f[-25:, 2], f[25:50, 2] + L)
        assert np.allclose(f[:, 0], f[0, 0]) and np.allclose(f[:, 1], f[0, 1])


def test_load_geometry_structures_and_sids():
    M, L = _two_fibers()
    g = sb.load_geometry(M, L, D=2.0, T2=80.0, D_box=3.0)
    assert g.nstructures == 2 and g.nsphere == 2 * (NZ + 50)
    assert np.array_equal(np.unique(g.spheres[4]), [0, 1])
    assert g.structure_D().tolist() == [2.0, 2.0, 3.0]
    assert g.structure_T2().tolist() == [80.0, 80.0, 80.0]


def test_generator_is_seeded_and_periodic():
    M1, L = sb.generate_beaded_axons(12, 30.0, seed=7)
    M2, _ = sb.generate_beaded_axons(12, 30.0, seed=7)
    M3, _ = sb.generate_beaded_axons(12, 30.0, seed=8)
    assert np.array_equal(M1, M2) and not np.array_equal(M1, M3)
    fibers = sb.split_matrix_to_list(M1, L)
    assert len(fibers) >= 12
    base = [f for f in fibers if np.all(np.abs(f[:, :2]) <= L / 2)]
    for f in base:
        assert f[0, 2] == -L / 2 and f[-1, 2] == L / 2
        assert np.all(np.diff(f[:, 2]) > 0)
    # every xy image is a copy of a fiber shifted by +-L along x and/or y
    cent = {(round(f[0, 0], 9), round(f[0, 1], 9)) for f in base}
    for f in fibers:
        x, y = f[0, 0], f[0, 1]
        assert any((round(x - dx, 9), round(y - dy, 9)) in cent
                   for dx in (-L, 0, L) for dy in (-L, 0, L))


def test_generator_packing_has_no_overlap():
    M, L = sb.generate_beaded_axons(30, 30.0, seed=3, gap=0.1)
    fibers = [f for f in sb.split_matrix_to_list(M, L) if np.all(np.abs(f[:, :2]) <= L / 2)]
    c = np.array([f[0, :2] for f in fibers])
    rmax = np.array([f[:, 3].max() for f in fibers])
    for i in range(len(c)):
        for j in range(i + 1, len(c)):
            d = c[i] - c[j]
            d -= L * np.round(d / L)
            assert np.hypot(*d) >= rmax[i] + rmax[j] + 0.1 - 1e-9


def test_generator_bead_law():
    M, L = sb.generate_beaded_axons(5, 30.0, seed=11)
    for f in sb.split_matrix_to_list(M, L)[:5]:
        r = f[:, 3]
        r0 = (r.max() + r.min()) / 2  # r0 (1 +- 0.3) extremes over >= one wavelength
        assert r.max() / r0 == pytest.approx(1.3, abs=0.05)
        assert r.min() > 0.25  # consecutive beads 0.5 um apart overlap


def test_geometry_isinside_rules():
    g = SimGeometry3D(10, 10, 10, 2.0, 100, 1)
    s = Structure3D(np.array([0.0]), np.array([0.0]), np.array([0.0]), np.array([1.0]), 2.0, 100, 1)
    g.add_structure(s)
    pts = np.array([[0.0, 0.99, 1.01, 5.0], [0, 0, 0, 0], [0, 0, 0, 0]])
    assert s.isinside(pts).tolist() == [True, True, False, False]
    assert g.isinside(pts).tolist() == [True, True, True, False]  # box: strict inequalities
    assert g.structures[-1] is g and g.structures[0] is s
