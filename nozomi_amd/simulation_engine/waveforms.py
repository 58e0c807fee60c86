"""Diffusion gradient waveforms sampled on the walk's time grid.

Mirrors simulation_toolkit/simulation_engine/waveforms.py:136-369: pulsed
(PGSE), cosine oscillating (OGSE) and apodized-cosine OGSE waveforms with the
same numpy sampling calls (so sample placement is identical), and the b-value
of a sampled waveform (equation 15 of doi:10.1002/nbm.1520, waveforms.py:147-153).

Units follow the reference: ``wave`` is in the reference's gradient unit, in
which ``wave / 1e4`` is taken as mT/um (waveforms.py:144,152; the per-step
phase coefficient of diffsim3d.py:254 uses the same factor), t and dt in ms,
b in ms/um^2 (1 ms/um^2 = 1000 s/mm^2).

The linear-programming lobe designers grad_lobe_by_area / grad_lobe_by_dur
(waveforms.py:8-133) have no caller in the reference and are not provided.
"""
import numpy as np

gamma = 267.513  # rad/ms/mT  (waveforms.py:4)
# gyromagnetic constants used by the reference's DWI driver for the per-step
# phase and for the wrap area (diffsim3d.py:238-239,254-255)
GAMMA_PHASE = 267.5
GAMMA_AREA = 267.7
UNIT_SCALE = 1e4  # wave / UNIT_SCALE = mT/um in the reference's convention

dt0 = 0.0001  # default time step (ms), waveforms.py:6


class DiffGradWaveform:
    """b-value helpers (waveforms.py:136-162)."""

    def calculate_bvalue(self, gmax=None):
        """b (ms/um^2) of gmax * wave.  The reference reads a non-existent
        ``self.gwave`` here (waveforms.py:144), which makes set_bvalues and
        set_gmax raise; this version uses ``self.wave``."""
        if gmax is None:
            gmax = self.gmax
        gwave_um = self.wave / 10e3
        return np.sum(np.cumsum(gmax * gamma * gwave_um) ** 2) * self.dt ** 3

    def calculate_bvalue_from_wave(self):
        """b (ms/um^2) of the sampled waveform (waveforms.py:147-153)."""
        gwave_um = self.wave / 10e3
        return np.sum(np.cumsum(gamma * gwave_um) ** 2) * self.dt ** 3

    def set_bvalues(self, b):
        b0 = self.calculate_bvalue(1)
        self.b = b
        self.gmax = np.sqrt(b / b0)

    def set_gmax(self, gmax):
        self.gmax = gmax
        self.b = self.calculate_bvalue(gmax)


class PGDiffWaveform(DiffGradWaveform):
    """Pulsed gradient: unit lobes +1 / -1 of duration little_delta separated by
    big_delta, centred on te/2 (waveforms.py:164-186).  Samples at
    t = dt, 2dt, ..., te; lobe membership uses strict inequalities."""

    def __init__(self, big_delta, little_delta, te, time_step=dt0):
        self.big_delta = big_delta
        self.little_delta = little_delta
        self.te = te
        self.dt = time_step
        self.generate_waveform()

    def generate_waveform(self):
        self.t = np.arange(self.dt, self.te + self.dt, self.dt)
        self.wave = np.zeros_like(self.t)
        first_lobe = (self.t > (self.te / 2 - self.big_delta / 2 - self.little_delta / 2)) \
            & (self.t < (self.te / 2 - self.big_delta / 2 + self.little_delta / 2))
        self.wave[first_lobe] = 1
        second_lobe = (self.t > (self.te / 2 + self.big_delta / 2 - self.little_delta / 2)) \
            & (self.t < (self.te / 2 + self.big_delta / 2 + self.little_delta / 2))
        self.wave[second_lobe] = -1


class CosineOGDiffWaveform(DiffGradWaveform):
    """Cosine OGSE with quarter-sine ramps in each half of te
    (waveforms.py:188-247).  As in the reference, the period is computed in ms
    from a frequency in Hz while the cosine phase uses 2*pi*f*t with t in ms."""

    def __init__(self, num_oscillations, frequency, gmax, te, time_step=dt0):
        self.num_oscillations = num_oscillations
        self.frequency = frequency
        self.gmax = gmax
        self.te = te
        self.dt = time_step
        self.generate_waveform()

    def generate_waveform(self):
        self.t = np.arange(self.dt, self.te + self.dt, self.dt)
        self.wave = np.zeros_like(self.t)
        period = 1.0 / self.frequency * 1000
        cos_dur = self.num_oscillations * period
        ramp = period / 8
        half = cos_dur + 2 * ramp
        for start, pol in ((self.te / 4 - half / 2, 1.0), (3 * self.te / 4 - half / 2, -1.0)):
            for i, t in enumerate(self.t):
                if start <= t < start + ramp:
                    self.wave[i] = pol * self.gmax * np.sin((t - start) / ramp * np.pi / 2)
                elif start + ramp <= t < start + ramp + cos_dur:
                    self.wave[i] = pol * self.gmax * np.cos(
                        2 * np.pi * self.frequency * (t - (start + ramp)))
                elif start + ramp + cos_dur <= t < start + 2 * ramp + cos_dur:
                    self.wave[i] = pol * self.gmax * np.cos(
                        (t - (start + ramp + cos_dur)) / ramp * np.pi / 2)


class ApodizedCosineOGSEWaveform(DiffGradWaveform):
    """Apodized cosine OGSE (Does et al. 2003): 1/4 sine at 2f, N cosine cycles
    at f, 1/4 sine at 2f; encoding lobe centred on te/4 and decoding lobe
    (negated) on 3te/4 (waveforms.py:250-369).  Samples start at t = 0."""

    def __init__(self, N_cycles, T_duration, te=None, gmax=1.0, time_step=dt0):
        self.N_cycles = N_cycles
        self.T_duration = T_duration
        self.gmax = gmax
        self.dt = time_step
        self.t_eff = T_duration / (4 * N_cycles)
        self.te = 4 * T_duration if te is None else te
        self.generate_waveform()

    def _timing(self):
        f_cosine = (0.25 + self.N_cycles) / self.T_duration
        f_apod = 2 * f_cosine
        return f_cosine, f_apod, 0.25 / f_apod, self.N_cycles / f_cosine

    def generate_waveform(self):
        self.t = np.arange(0, self.te + self.dt, self.dt)
        self.wave = np.zeros_like(self.t)
        _, _, apod, cos_dur = self._timing()
        self._generate_single_waveform(self.te / 4 - self.T_duration / 2, 1.0, apod, cos_dur)
        self._generate_single_waveform(3 * self.te / 4 - self.T_duration / 2, -1.0, apod, cos_dur)

    def _generate_single_waveform(self, t_start, polarity, apod_duration, cosine_duration):
        for i, t in enumerate(self.t):
            t_rel = t - t_start
            if 0 <= t_rel < apod_duration:
                self.wave[i] = polarity * self.gmax * np.sin((t_rel / apod_duration) * (np.pi / 2))
            elif apod_duration <= t_rel < apod_duration + cosine_duration:
                progress = (t_rel - apod_duration) / cosine_duration
                self.wave[i] = polarity * self.gmax * np.cos(2 * np.pi * self.N_cycles * progress)
            elif apod_duration + cosine_duration <= t_rel < 2 * apod_duration + cosine_duration:
                t_final = t_rel - (apod_duration + cosine_duration)
                self.wave[i] = polarity * self.gmax * np.cos((t_final / apod_duration) * (np.pi / 2))

    def get_effective_diffusion_time(self):
        return self.t_eff

    def get_waveform_info(self):
        f_cosine, f_apod, apod, cos_dur = self._timing()
        return {
            'N_cycles': self.N_cycles,
            'T_duration_ms': self.T_duration,
            'effective_diffusion_time_ms': self.t_eff,
            'total_echo_time_ms': self.te,
            'apodization_duration_ms': apod,
            'cosine_duration_ms': cos_dur,
            'cosine_frequency_Hz': f_cosine * 1000,
            'apodization_frequency_Hz': f_apod * 1000,
            'frequency_ratio': f_apod / f_cosine,
        }


def phase_coefficients(wave, dt):
    """Per-step phase coefficient G and wrap area G_area of a sampled waveform
    as the reference's DWI driver forms them (diffsim3d.py:250-255):
    G_area = trapezoid cumulative integral of wave (dx = 1, initial 0) * dt;
    G = wave * dt * 267.5 / 1e4, area * 267.7 / 1e4.  Both are rounded to
    float32 like the reference's kernel arguments (diffsim3d.py:224)."""
    wave = np.asarray(wave, dtype=np.float64)
    area = np.zeros_like(wave)
    if wave.size > 1:
        area[1:] = np.cumsum((wave[1:] + wave[:-1]) * 0.5)
    area = area * dt
    g = (wave * dt * GAMMA_PHASE / UNIT_SCALE).astype(np.float32)
    ga = (area * GAMMA_AREA / UNIT_SCALE).astype(np.float32)
    return g, ga


def bvalue_of_coefficients(g, dt):
    """b (ms/um^2) seen by the walker for per-step phase coefficients g
    (rad/um per step): b = dt * sum_n (sum_{k<=n} g_k)^2."""
    q = np.cumsum(np.asarray(g, dtype=np.float64))
    return float(dt * np.sum(q * q))
