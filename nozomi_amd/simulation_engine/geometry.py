"""Substrate model: structures as unions of spheres inside a periodic box.

Mirrors simulation_toolkit/simulation_engine/geometry.py: a structure is a
union of spheres; SimGeometry3D keeps every sphere in ``spheres[5, S]`` =
(cx, cy, cz, r, sid) with sid = insertion order = precedence (lower wins,
sim3d_kernel.cu:53-54), and the box itself as the last entry of
``structures`` (geometry.py:37-67).
"""
import numpy as np


class Structure3D:
    """A union of spheres with its own D (um^2/ms), T2 (ms) and density rho
    (geometry.py:8-35)."""

    def __init__(self, x, y, z, r, D, T2, rho):
        x, y, z, r = (np.asarray(v, dtype=np.float64) for v in (x, y, z, r))
        self.nspheres = x.size
        self.x = np.reshape(x, [self.nspheres, 1])
        self.y = np.reshape(y, [self.nspheres, 1])
        self.z = np.reshape(z, [self.nspheres, 1])
        self.r = np.reshape(r, [self.nspheres, 1])
        self.D = D
        self.T2 = T2
        self.rho = rho
        self.tol = 1e-4

    def isinside(self, rnd):
        """True for every column of rnd [3, N] inside some sphere (geometry.py:26-35)."""
        rnd = np.asarray(rnd)
        inside_any = ((rnd[0, :] - self.x) ** 2 + (rnd[1, :] - self.y) ** 2 +
                      (rnd[2, :] - self.z) ** 2 - self.r ** 2 < 0)
        return inside_any.any(axis=0)


class SimGeometry3D:
    """Periodic box [-L/2, L/2]^3 holding structures (geometry.py:37-71)."""

    def __init__(self, Lx, Ly, Lz, D, T2, rho):
        self.Lx = Lx
        self.Ly = Ly
        self.Lz = Lz
        self.nsphere = 0
        self.spheres = np.zeros([5, self.nsphere])
        self.nstructures = 0
        # the last item is always the box itself
        self.structures = [self]
        self.D = D
        self.T2 = T2
        self.rho = rho

    def add_structure(self, struct):
        """Append a structure; its spheres get sid = current structure count
        (geometry.py:54-67)."""
        self.structures[-1] = struct
        self.structures.append(self)
        newsphere = np.zeros([5, struct.nspheres])
        newsphere[0, :] = struct.x.T
        newsphere[1, :] = struct.y.T
        newsphere[2, :] = struct.z.T
        newsphere[3, :] = struct.r.T
        newsphere[4, :] = self.nstructures
        self.spheres = np.concatenate([self.spheres, newsphere], axis=1)
        self.nstructures += 1
        self.nsphere += struct.nspheres

    def isinside(self, rnd):
        """Strictly inside the box (geometry.py:69-71)."""
        rnd = np.asarray(rnd)
        return ((rnd[0, :] > -self.Lx / 2) & (rnd[1, :] > -self.Ly / 2) & (rnd[2, :] > -self.Lz / 2) &
                (rnd[0, :] < self.Lx / 2) & (rnd[1, :] < self.Ly / 2) & (rnd[2, :] < self.Lz / 2))

    def structure_D(self):
        """Per-structure D with the box last (diffsim3d.py:95-98)."""
        return np.array([s.D for s in self.structures], dtype=np.float64)

    def structure_T2(self):
        """Per-structure T2 with the box last (diffsim3d.py:100-103)."""
        return np.array([s.T2 for s in self.structures], dtype=np.float64)

    def device_substrate(self, cell_size=0.0, margin=0.0, device=None):
        """Upload the spheres and build the cell grid (nz_substrate_create)."""
        from ..runtime import DeviceSubstrate
        return DeviceSubstrate(self.spheres, self.nstructures, self.structure_D(),
                               self.structure_T2(), (self.Lx, self.Ly, self.Lz),
                               cell_size=cell_size, margin=margin, device=device)

    def gpu_seed(self, gpu_seed_kernel=None, nspin=None, structIdxs=None, seed=None,
                 substrate=None):
        """Seed nspin spins uniformly in the structures structIdxs (None = every
        structure and the box) on the GPU; returns a float32 numpy [3, nspin]
        (geometry.py:73-174).  The first argument exists for call compatibility
        with the reference, which passes its compiled PyCUDA kernel there; it
        is ignored.  seed: Philox key (None draws one from numpy)."""
        if nspin is None:
            raise TypeError("gpu_seed needs nspin")
        if structIdxs is None:
            structIdxs = np.arange(self.nstructures + 1, dtype=np.int32)
        if seed is None:
            seed = int(np.random.randint(np.iinfo(np.int32).max))
        sub = substrate if substrate is not None else self.device_substrate()
        pos = sub.seed(int(nspin), targets=np.asarray(structIdxs, dtype=np.int32), seed=seed)
        return pos.cpu().numpy()
