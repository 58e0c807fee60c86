"""Walker host API: DiffSim3d (narrow-pulse walk) and DwiSim3d (gradient
phases), backed by libnozomi_hip.so.

Mirrors simulation_toolkit/simulation_engine/diffsim3d.py.  The reference
%-formats and JIT-compiles a CUDA template per geometry and launches one
kernel per time step; here the substrate is uploaded once (cell grid built by
nz_substrate_create), the spin state stays resident, and a call advances any
number of steps in one launch (``step`` keeps the reference's one-step
semantics; ``run`` and ``DwiSim3d.simulate*`` advance many steps per launch).

Device arrays are exposed as DeviceArray views (``.get()`` returns numpy) with
the reference's shapes: spins_d / spins0_d [3, N], sig_d [1, N], phase_d
[ndiffdir * N].
"""
import numpy as np
import torch

from .. import runtime as rt
from ..runtime import DeviceArray
from . import waveforms as wf


class DiffSim3d:
    """Simulates diffusion within a 3D geometry of structures (diffsim3d.py:8-226)."""

    def __init__(self, geom, nspins, diffdir=None, device=None, cell_size=None, margin=None,
                 sort_interval=None, counters=False):
        self.device = rt._require_gpu(device)
        self.cell_size = cell_size
        self.margin = margin
        self.sort_interval = sort_interval
        self.counters = counters
        self.sub = None
        self.walker = None
        self.seed = None
        self.set_spins(nspins)
        self.set_geometry(geom)
        self.set_diffusion_directions(diffdir)
        self.set_block_grid(nspins)
        self.nsegx = self.nsegy = self.nsegz = 1
        self.issetup = False

    # -- configuration (diffsim3d.py:65-104,176-178) -------------------------
    def set_diffusion_directions(self, diffdir=None):
        if diffdir is None:
            diffdir = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]])
        if np.isscalar(diffdir):
            raise TypeError("diffdir cannot be a scalar (so far)")
        diffdir = np.asarray(diffdir)
        if (diffdir.ndim != 2) or (diffdir.shape[0] != 3):
            raise TypeError("diffdir must be a numpy array of size 3xN")
        self.ndiffdir = diffdir.shape[1]
        self.diffdir_host = diffdir.astype(np.float32)
        self.diffdir = torch.from_numpy(np.ascontiguousarray(self.diffdir_host)).to(self.device)
        self.issetup = False

    def set_spins(self, nspins):
        nspins = int(nspins)
        if nspins <= 0:
            raise ValueError("nspins must be positive")
        self.nspins = nspins
        self.issetup = False

    def set_geometry(self, geom):
        self.geom = geom
        self.sub = None
        self.issetup = False

    def set_block_grid(self, nspins):
        """Launch shape is chosen by the library (one spin per lane, 64-lane
        workgroups for the walk); kept for API compatibility
        (diffsim3d.py:176-178)."""
        self.nblock = 64
        self.ngrid = (int(nspins) + 63) // 64

    def set_segments(self, nsegx=1, nsegy=1, nsegz=1):
        """The reference builds padded per-segment sphere lists here
        (diffsim3d.py:107-174).  The cell grid of this implementation is built
        when the substrate is uploaded and its lists are complete for any jump,
        so the segment counts do not change results; they are recorded, and
        used as the cell size only when DiffSim3d(cell_size=...) is not given
        and the segments are finer than the automatic choice."""
        self.nsegx, self.nsegy, self.nsegz = int(nsegx), int(nsegy), int(nsegz)
        self.sub = None
        self.issetup = False

    def _cell_size(self):
        if self.cell_size:
            return float(self.cell_size)
        auto = 2.0
        L = min(self.geom.Lx, self.geom.Ly, self.geom.Lz)
        seg = L / max(self.nsegx, self.nsegy, self.nsegz)
        return max(min(auto, seg), 0.5) if max(self.nsegx, self.nsegy, self.nsegz) > 1 else auto

    def _substrate(self):
        if self.sub is None:
            self.sub = self.geom.device_substrate(cell_size=self._cell_size(),
                                                  margin=self.margin or 0.0, device=self.device)
        return self.sub

    # -- set-up (diffsim3d.py:35-63) -----------------------------------------
    def setup(self, structures=None, initstates_int=None, positions=None):
        """Seed the spins (in ``structures``; None = every structure and the
        box) and build the resident state.  initstates_int is the Philox key
        of the walk (None draws one from numpy, as the reference draws its
        cuRAND seed, diffsim3d.py:56-58).  positions: optional [3, N] start
        positions (numpy or tensor) instead of seeding."""
        sub = self._substrate()
        if initstates_int is None:
            initstates_int = int(np.random.randint(np.iinfo(np.int32).max, dtype=np.int32))
        self.seed = int(initstates_int)
        self.walker = rt.Walker(sub, self.nspins, counters=self.counters,
                                sort_interval=self.sort_interval)
        if positions is not None:
            p = positions if torch.is_tensor(positions) else torch.from_numpy(
                np.ascontiguousarray(positions, dtype=np.float32))
            self.walker.set_positions(p.to(self.device, torch.float32))
        else:
            if structures is None:
                structures = np.arange(self.geom.nstructures + 1, dtype=np.int32)
            # the seeding stream is keyed apart from the walk stream
            self.walker.seed(targets=structures, seed=self.seed ^ 0x5EED5EED5EED)
        f32 = torch.float32
        self._pos = torch.empty((3, self.nspins), dtype=f32, device=self.device)
        self._pos0 = torch.empty((3, self.nspins), dtype=f32, device=self.device)
        self._sig = torch.empty((1, self.nspins), dtype=f32, device=self.device)
        self._mom = torch.empty((3, self.nspins), dtype=f32, device=self.device)
        self._fresh = False
        # which reference kernel the phase buffer belongs to: "axes" after
        # dwi_step / simulate (randomWalk3d_phase writes x, y, z phases),
        # "directions" after dwi_multidirections_step (one row per diffdir)
        self._phase_kind = "directions"
        self.issetup = True
        self.spins = self.spins_d.get()

    def _views(self):
        if not self._fresh:
            self.walker.read(self._pos, self._pos0, self._sig, self._mom)
            self._fresh = True

    # -- device views ----------------------------------------------------------
    @property
    def spins_d(self):
        self._views()
        return DeviceArray(self._pos)

    @property
    def spins0_d(self):
        self._views()
        return DeviceArray(self._pos0)

    @property
    def sig_d(self):
        self._views()
        return DeviceArray(self._sig)

    @property
    def mom_d(self):
        """[3, N] phase moment sum_t G_t x_t (with wrap corrections)."""
        self._views()
        return DeviceArray(self._mom)

    @property
    def phase_d(self):
        """The reference's phase buffer, ndiffdir * N values (diffsim3d.py:45).
        After dwi_multidirections_step: the phase of every direction of
        diffdir, the projection of the moment (sim3d_kernel.cu:880-942).  After
        dwi_step: randomWalk3d_phase's x, y, z phases in the first three rows
        (sim3d_kernel.cu:674-715), the rows beyond them zero as allocated."""
        self._views()
        if self._phase_kind == "axes":
            ph = torch.zeros((self.ndiffdir, self.nspins), dtype=torch.float32, device=self.device)
            k = min(3, self.ndiffdir)
            ph[:k] = self._mom[:k]
        else:
            ph = rt.project_phase(self._mom, self.diffdir)
        return DeviceArray(ph.reshape(-1))

    # -- stepping (diffsim3d.py:180-226) ---------------------------------------
    def _ensure(self):
        if not self.issetup:
            self.setup()

    def _walk(self, nsteps, dt, **kw):
        self._ensure()
        self.walker.walk(self.seed, nsteps, dt, **kw)
        self._fresh = False

    def step(self, dt):
        """One time step with the T2 weight (randomWalk3d)."""
        self._walk(1, dt, t2=True)

    def run(self, nsteps, dt, stepmom=None, t2=True):
        """nsteps time steps in one call; stepmom: optional float64 tensor
        [nsteps, 6] receiving per-step displacement moment sums."""
        self._walk(int(nsteps), dt, stepmom=stepmom, t2=t2)

    def _phase_step(self, gwave_dt, G, G_area):
        g = torch.tensor([np.float32(G)], dtype=torch.float32, device=self.device)
        ga = torch.tensor([np.float32(G_area)], dtype=torch.float32, device=self.device)
        self._walk(1, gwave_dt, g=g, garea=ga, t2=False)

    def dwi_step(self, gwave_dt, G, G_area_at_each_time_step):
        """One step accumulating x.G per axis (randomWalk3d_phase); phase_d
        then holds the axis phases."""
        self._phase_step(gwave_dt, G, G_area_at_each_time_step)
        self._phase_kind = "axes"

    def dwi_multidirections_step(self, gwave_dt, G, G_area_at_each_time_step):
        """One step accumulating (x.u_d)G per direction (randomWalk3d_phase_multidirr);
        phase_d then holds the direction phases."""
        self._phase_step(gwave_dt, G, G_area_at_each_time_step)
        self._phase_kind = "directions"

    def calculate_diffusion_coefficients_and_kurtoses(self, step_idx, current_time,
                                                      dx_array, dy_array, dz_array,
                                                      Kx2_array, Ky2_array, Kz2_array,
                                                      Kx4_array, Ky4_array, Kz4_array):
        """Add this step's D = <dx^2>/(2t), K2 = <dx^2>, K4 = <dx^4> per axis
        into element step_idx of the nine arrays (diffsim3d.py:180-197,
        sim3d_kernel.cu:162-238).  The arrays are DeviceArray / tensors /
        numpy arrays."""
        self._ensure()
        s = self.walker.displacement_moments()
        N = float(self.nspins)
        vals = torch.stack([s[0] / (2.0 * current_time * N), s[1] / (2.0 * current_time * N),
                            s[2] / (2.0 * current_time * N), s[0] / N, s[1] / N, s[2] / N,
                            s[3] / N, s[4] / N, s[5] / N])
        arrays = (dx_array, dy_array, dz_array, Kx2_array, Ky2_array, Kz2_array,
                  Kx4_array, Ky4_array, Kz4_array)
        for a, v in zip(arrays, vals):
            t = a.tensor if isinstance(a, DeviceArray) else a
            if torch.is_tensor(t):  # on any device (a host tensor takes the value over)
                t[step_idx] += v.to(t.device, t.dtype)
            else:
                t[step_idx] += float(v)


class DwiSim3d(DiffSim3d):
    """DWI phases under a gradient waveform (diffsim3d.py:228-261).  The whole
    waveform is walked in one call."""

    def _coefficients(self, gwave):
        g, ga = wf.phase_coefficients(gwave.wave, gwave.dt)
        return (torch.from_numpy(g).to(self.device), torch.from_numpy(ga).to(self.device))

    def _simulate(self, gwave, structures, initstates, kind):
        if not self.issetup:
            self.setup(structures=structures, initstates_int=initstates)
        g, ga = self._coefficients(gwave)
        self._walk(len(gwave.wave), gwave.dt, g=g, garea=ga, t2=False)
        self._phase_kind = kind

    def simulate(self, gwave, structures=None, initstates=None):
        """Phases along x, y, z (the dwi_step kernel), returned as the
        reference returns its phase buffer: numpy [ndiffdir, N], the x, y, z
        phases in the first three rows (diffsim3d.py:230-244; with the default
        diffdir that is [3, N])."""
        self._simulate(gwave, structures, initstates, "axes")
        return self.phase_d.get().reshape(self.ndiffdir, self.nspins)

    def simulate_multi_directions(self, gwave, structures=None, initstates=None):
        """Phases along every direction of diffdir: numpy [ndiffdir, N]
        (diffsim3d.py:246-261)."""
        self._simulate(gwave, structures, initstates, "directions")
        return self.phase_d.get().reshape(self.ndiffdir, self.nspins)

    def signals(self, scales=None):
        """Complex signal per direction, sum_i exp(i*phase)/N, computed in FP64
        on the device from the current moments; scales multiplies each
        direction's phase (b-shells of one waveform shape)."""
        self._views()
        sc = None if scales is None else torch.as_tensor(
            np.asarray(scales, dtype=np.float32)).to(self.device)
        s = rt.signal_sums(self._mom, self.diffdir, sc).cpu().numpy()
        return (s[:, 0] + 1j * s[:, 1]) / self.nspins
