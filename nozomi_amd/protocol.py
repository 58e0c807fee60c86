"""Acquisition protocols for the DWI walker: gradient directions and PGSE
shells sampled on the walk's time grid.

A shell set of one waveform shape is walked once: every (shell, direction)
phase is scale_shell * u_dir . m, where m is the per-spin moment the walker
accumulates (see DwiSim3d).  The b-value of a sampled waveform is that of the
walker's discrete phase, b = dt * sum_n (sum_{k<=n} g_k)^2, so free diffusion
gives exactly exp(-b D) in expectation.
"""
import numpy as np

from .simulation_engine import waveforms as wf


def fibonacci_hemisphere(n):
    """n unit vectors on the z >= 0 hemisphere by the Fibonacci lattice.
    Returns [3, n]."""
    k = np.arange(n) + 0.5
    z = 1.0 - k / n                      # (0, 1]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = np.pi * (3.0 - np.sqrt(5.0)) * np.arange(n)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z])


def icosahedron_6():
    """The 6 non-antipodal vertex directions of a regular icosahedron, [3, 6]."""
    t = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([[0, 1, t], [0, -1, t], [1, t, 0], [-1, t, 0], [t, 0, 1], [-t, 0, 1]], float)
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).T


class WaveformShells:
    """One gradient waveform shape (unit amplitude samples ``wave`` on a grid
    of step dt, ms) played at several b-values (ms/um^2) along directions
    [3, ndir].

    wave: the waveform scaled so that its walk phase has b = b_values[0];
    g, garea: float32 per-step coefficients (diffsim3d.py:250-255);
    dirs [3, W] and scales [W] list every (shell, direction) signal, shell
    major; b_per_signal [W] is the b-value of each signal's discrete phase."""

    def __init__(self, wave, dt, b_values, dirs, t=None):
        self.dt = float(dt)
        self.b_values = np.atleast_1d(np.asarray(b_values, dtype=np.float64))
        if self.b_values.size == 0 or np.any(self.b_values <= 0):
            raise ValueError("b_values must be positive")
        unit = np.asarray(wave, dtype=np.float64)
        g_unit, _ = wf.phase_coefficients(unit, self.dt)
        b_unit = wf.bvalue_of_coefficients(g_unit.astype(np.float64), self.dt)
        if not b_unit > 0:
            raise ValueError("the waveform has no diffusion weighting")
        self.amplitude = np.sqrt(self.b_values[0] / b_unit)
        self.wave = unit * self.amplitude
        self.t = t
        self.g, self.garea = wf.phase_coefficients(self.wave, self.dt)
        self.b_walk = wf.bvalue_of_coefficients(self.g.astype(np.float64), self.dt)
        d = np.asarray(dirs, dtype=np.float64)
        if d.ndim == 1:
            d = d.reshape(3, 1)
        if d.shape[0] != 3:
            raise ValueError("dirs must be [3, ndir]")
        d = d / np.linalg.norm(d, axis=0, keepdims=True)
        self.ndir = d.shape[1]
        self.unit_dirs = d
        self.dirs = np.tile(d, (1, self.b_values.size))
        self.scales = np.repeat(np.sqrt(self.b_values / self.b_values[0]), self.ndir)
        self.b_per_signal = self.b_walk * self.scales ** 2

    @property
    def nsteps(self):
        return self.g.size

    @property
    def nsignals(self):
        return self.dirs.shape[1]


class PGSEShells(WaveformShells):
    """PGSE with little_delta, big_delta (ms) on a grid of step dt (ms), te =
    big_delta + little_delta unless given, at b-values (ms/um^2) and
    directions [3, ndir] (waveforms.py:164-186)."""

    def __init__(self, little_delta, big_delta, dt, b_values, dirs, te=None):
        self.little_delta = float(little_delta)
        self.big_delta = float(big_delta)
        self.te = self.big_delta + self.little_delta if te is None else float(te)
        base = wf.PGDiffWaveform(self.big_delta, self.little_delta, self.te, float(dt))
        super().__init__(base.wave, dt, b_values, dirs, t=base.t)


class OGSEShells(WaveformShells):
    """Apodized cosine OGSE with N_cycles cycles per lobe of duration
    T_duration (ms) in an echo time te (waveforms.py:250-343)."""

    def __init__(self, N_cycles, T_duration, te, dt, b_values, dirs):
        self.N_cycles = int(N_cycles)
        self.T_duration = float(T_duration)
        self.te = float(te)
        base = wf.ApodizedCosineOGSEWaveform(self.N_cycles, self.T_duration, te=self.te,
                                             time_step=float(dt))
        self.info = base.get_waveform_info()
        super().__init__(base.wave, dt, b_values, dirs, t=base.t)
