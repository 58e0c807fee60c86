"""Substrates for the walker.

1. The reference's substrate feed: the pickle [Tensor(M, 5) (x, y, z, r, fid), L]
   written by its generator, split into fibers at z == L/2
   (utils/common_utils.py:17-55), z-padded with 25 image spheres per end
   (utils/adjust_geometry.py:3-36) and turned into one Structure3D per fiber
   (cli/simulation_main.py:37-54).
2. A seeded synthetic generator of beaded axons parallel to z in the same
   matrix format, for benchmark and test substrates: xy centres by random
   sequential packing with periodic distances, chains of spheres every 0.5 um
   with r(z) = r0 (1 + 0.3 sin(2 pi z / lambda + phi)), r0 ~ Gamma(4, 0.25 um),
   lambda ~ U(3, 8) um, phi ~ U(0, 2 pi), and xy periodic images added the way
   the reference generator adds them (substrate_generator/
   geometric_optimization.py:256-300).
"""
import pickle

import numpy as np

from .simulation_engine.geometry import SimGeometry3D, Structure3D


# ------------------------------------------------------------ reference feed
def import_array_geometry_full_path(file_name):
    """Unpickle [optimized_fibers, L] (common_utils.py:17-19).  Only for
    substrate files you trust: pickle can execute code."""
    with open(file_name, "rb") as f:
        fibers, L = pickle.load(f)
    if hasattr(fibers, "cpu"):
        fibers = fibers.cpu().numpy()
    if hasattr(L, "item"):
        L = L.item()
    return np.asarray(fibers, dtype=np.float64), float(L)


def save_data_array_to_pickle(file_name, optimized_fibers, L):
    """Write the reference substrate format (common_utils.py:21-25)."""
    with open(file_name, "wb") as f:
        pickle.dump([optimized_fibers, L], f, protocol=pickle.HIGHEST_PROTOCOL)


def split_matrix_to_list(A, box_length):
    """Split the flat sphere matrix into fibers: a fiber ends at the row whose
    z equals L/2 exactly (common_utils.py:41-55)."""
    out = []
    start = 0
    for i in range(len(A)):
        if A[i, 2] == box_length / 2:
            out.append(np.array(A[start:i + 1]))
            start = i + 1
    return out


def create_pad(end_fiber, start_fiber, Lz, pad=50):
    """Prepend the last pad/2 spheres of end_fiber shifted by -Lz and append
    its first pad/2 shifted by +Lz (adjust_geometry.py:22-36)."""
    h = pad // 2
    out = np.zeros((start_fiber.shape[0] + pad, start_fiber.shape[-1]))
    out[:, 0] = np.concatenate((end_fiber[-h:, 0], start_fiber[:, 0], end_fiber[:h, 0]))
    out[:, 1] = np.concatenate((end_fiber[-h:, 1], start_fiber[:, 1], end_fiber[:h, 1]))
    out[:, 2] = np.concatenate((end_fiber[-h:, 2] - Lz, start_fiber[:, 2], end_fiber[:h, 2] + Lz))
    out[:, 3] = np.concatenate((end_fiber[-h:, 3], start_fiber[:, 3], end_fiber[:h, 3]))
    out[:, 4] = np.concatenate((start_fiber[-h:, 4], start_fiber[:, 4], start_fiber[:h, 4]))
    return out


def createBoundaryZ(fiberlist_xyz_r_fid, Lz):
    """z-periodic padding: a fiber whose first (x, y) matches the last (x, y)
    of some fiber is padded from that fiber (adjust_geometry.py:3-20)."""
    for f in fiberlist_xyz_r_fid:
        f[:, -1] = f[:, -1].astype(int)
    ends = {}
    for f in fiberlist_xyz_r_fid:
        ends[(f[-1][0], f[-1][1])] = f
    out = []
    for f in fiberlist_xyz_r_fid:
        key = (f[0][0], f[0][1])
        out.append(create_pad(ends[key], f, Lz) if key in ends else f)
    return out


def build_geometry(fibers, L, D=2.0, T2=100.0, rho=1.0, D_box=None, T2_box=None):
    """SimGeometry3D with one Structure3D per fiber (simulation_main.py:41-54).
    D_box / T2_box: parameters of the extra-axonal space (default D / T2)."""
    sg = SimGeometry3D(L, L, L, D if D_box is None else D_box, T2 if T2_box is None else T2_box, rho)
    for f in fibers:
        sg.add_structure(Structure3D(f[:, 0], f[:, 1], f[:, 2], f[:, 3], D, T2, rho))
    return sg


def load_geometry(matrix, L, **kw):
    """Reference pipeline: matrix -> fibers -> z padding -> SimGeometry3D."""
    fibers = split_matrix_to_list(np.asarray(matrix, dtype=np.float64), L)
    fibers = createBoundaryZ(fibers, L)
    return build_geometry(fibers, L, **kw)


# ---------------------------------------------------------------- generator
def _rsa_centres(rng, radii, L, gap, max_tries=200000):
    """Random sequential packing of discs (radii) in the periodic square
    [-L/2, L/2]^2, largest first."""
    order = np.argsort(-radii)
    xy = np.zeros((radii.size, 2))
    placed = []
    for i in order:
        for _ in range(max_tries):
            c = (rng.random(2) - 0.5) * L
            ok = True
            for j in placed:
                d = c - xy[j]
                d -= L * np.round(d / L)
                if d @ d < (radii[i] + radii[j] + gap) ** 2:
                    ok = False
                    break
            if ok:
                xy[i] = c
                placed.append(i)
                break
        else:
            raise RuntimeError("random sequential packing failed: lower the axon count or radii")
    return xy


def wrap_pbc_images(M, L, tol=None):
    """Append xy periodic images of whole fibers near a face, as the reference
    generator does (geometric_optimization.py:256-300): fibers with a sphere
    within tol of x = -L/2 get a copy at x + L, within tol of x = +L/2 a copy
    at x - L; then the same in y over the result.  tol defaults to L/5 (:167)."""
    if tol is None:
        tol = L / 5
    x, y, z, r, fid = (M[:, k].copy() for k in range(5))

    def images(coord, shift_axis):
        nonlocal x, y, z, r, fid
        c = x if shift_axis == 0 else y
        lo = np.isin(fid, fid[c - r < -L / 2 + tol])
        hi = np.isin(fid, fid[L - c - r < L / 2 + tol])
        idx_lo = np.nonzero(lo)[0]
        xs = [x, x[idx_lo] + (L if shift_axis == 0 else 0)]
        ys = [y, y[idx_lo] + (L if shift_axis == 1 else 0)]
        zs, rs, fs = [z, z[idx_lo]], [r, r[idx_lo]], [fid, fid[idx_lo]]
        x, y, z, r, fid = (np.concatenate(v) for v in (xs, ys, zs, rs, fs))
        idx_hi = np.nonzero(np.concatenate([hi, np.zeros(idx_lo.size, bool)]))[0]
        x = np.concatenate([x, x[idx_hi] - (L if shift_axis == 0 else 0)])
        y = np.concatenate([y, y[idx_hi] - (L if shift_axis == 1 else 0)])
        z = np.concatenate([z, z[idx_hi]])
        r = np.concatenate([r, r[idx_hi]])
        fid = np.concatenate([fid, fid[idx_hi]])

    images(x, 0)
    images(y, 1)
    return np.stack([x, y, z, r, fid], axis=1)


def generate_beaded_axons(n_axons, L, seed, spacing=0.5, r0_shape=4.0, r0_scale=0.25,
                          bead_amp=0.3, lam_range=(3.0, 8.0), gap=0.1, r0_min=None):
    """Seeded beaded-axon substrate in the reference matrix format.

    Returns (matrix [M, 5] = (x, y, z, r, fid), L).  Every fiber runs from
    z = -L/2 to z = +L/2 inclusive (the fiber-end rule of
    split_matrix_to_list); xy images follow.  r0 is drawn from
    Gamma(r0_shape, r0_scale) truncated below at r0_min (default: the value
    that keeps the thinnest bead overlapping its neighbours,
    spacing / (2 (1 - bead_amp)) * 1.05)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if r0_min is None:
        r0_min = spacing / (2.0 * (1.0 - bead_amp)) * 1.05
    r0 = np.empty(n_axons)
    for i in range(n_axons):
        v = rng.gamma(r0_shape, r0_scale)
        while v < r0_min:
            v = rng.gamma(r0_shape, r0_scale)
        r0[i] = v
    lam = rng.uniform(lam_range[0], lam_range[1], n_axons)
    phi = rng.uniform(0.0, 2 * np.pi, n_axons)
    rmax = r0 * (1 + bead_amp)
    xy = _rsa_centres(rng, rmax, L, gap)
    nz = int(round(L / spacing))
    z = -L / 2 + spacing * np.arange(nz + 1)
    z[-1] = L / 2  # exact end marker
    rows = []
    for i in range(n_axons):
        r = r0[i] * (1 + bead_amp * np.sin(2 * np.pi * z / lam[i] + phi[i]))
        rows.append(np.stack([np.full(z.size, xy[i, 0]), np.full(z.size, xy[i, 1]), z, r,
                              np.full(z.size, float(i))], axis=1))
    M = np.concatenate(rows, axis=0)
    return wrap_pbc_images(M, L), float(L)


def beaded_geometry(n_axons, L, seed, D=2.0, T2=100.0, D_box=None, **kw):
    """Generator + reference feed pipeline in one call."""
    M, L = generate_beaded_axons(n_axons, L, seed, **kw)
    return load_geometry(M, L, D=D, T2=T2, D_box=D_box)


def single_sphere_geometry(L=10.0, r=1.0, D=3.0, T2=100.0):
    """One sphere of radius r at the centre of the box (test_02...py:17-35)."""
    sg = SimGeometry3D(L, L, L, D, T2, 1)
    sg.add_structure(Structure3D(np.array([0.0]), np.array([0.0]), np.array([0.0]),
                                 np.array([r]), D, T2, 1))
    return sg


def cylinder_geometry(a=0.5, L=20.0, D=2.0, T2=100.0, spacing=None):
    """A straight cylinder of radius a along z, approximated by a dense
    sphere chain that spans the box plus 25 image spheres per end (the
    calibration study's single-cylinder substrate)."""
    if spacing is None:
        spacing = a / 10.0
    n = int(round(L / spacing))
    z = -L / 2 - 25 * spacing + spacing * np.arange(n + 51)
    sg = SimGeometry3D(L, L, L, D, T2, 1)
    sg.add_structure(Structure3D(np.zeros_like(z), np.zeros_like(z), z, np.full_like(z, a), D, T2, 1))
    return sg
