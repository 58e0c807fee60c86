"""Gradient waveforms and the b-value of the walker's discrete phase
(waveforms.py:136-369, diffsim3d.py:250-255), pinned by closed forms."""
import numpy as np
import pytest

from nozomi_amd.protocol import PGSEShells, fibonacci_hemisphere, icosahedron_6
from nozomi_amd.simulation_engine import waveforms as wf


def test_pgse_lobes_and_balance():
    w = wf.PGDiffWaveform(big_delta=30.0, little_delta=10.0, te=40.0, time_step=0.02)
    assert w.wave.size == 2000
    assert (w.wave == 1).sum() == (w.wave == -1).sum()
    assert w.wave.sum() == 0
    # strict inequalities: a lobe of duration 10 ms has 499 samples at 20 us
    assert (w.wave == 1).sum() == 499
    # lobes start at te/2 -/+ Delta/2 - delta/2
    t_first = w.t[np.argmax(w.wave == 1)]
    assert t_first == pytest.approx(0.02)
    t_second = w.t[np.argmax(w.wave == -1)]
    assert t_second == pytest.approx(30.02, abs=0.021)


def test_bvalue_closed_form_rectangular_pgse():
    # b = q^2 (Delta - delta/3) for rectangular lobes, q = gamma G delta, with
    # the sampled lobes' own duration and separation; the rectangle-rule sum
    # converges to it as dt -> 0
    D, d = 30.0, 10.0
    for dt, tol in ((0.01, 1e-3), (0.002, 2e-4)):
        w = wf.PGDiffWaveform(D, d, D + d, dt)
        b = w.calculate_bvalue_from_wave()
        n = int((w.wave == 1).sum())
        d_eff = n * dt
        D_eff = (np.argmax(w.wave == -1) - np.argmax(w.wave == 1)) * dt
        q = wf.gamma * (1.0 / 10e3) * d_eff
        assert b == pytest.approx(q * q * (D_eff - d_eff / 3), rel=tol)


def test_set_bvalues_uses_wave():
    w = wf.PGDiffWaveform(30.0, 10.0, 40.0, 0.02)
    w.set_bvalues(1.0)
    assert w.calculate_bvalue(w.gmax) == pytest.approx(1.0)


def test_phase_coefficients_match_reference_formulas():
    w = wf.PGDiffWaveform(30.0, 10.0, 40.0, 0.02)
    g, ga = wf.phase_coefficients(w.wave, w.dt)
    assert g.dtype == np.float32 and ga.dtype == np.float32
    from scipy.integrate import cumulative_trapezoid
    area = cumulative_trapezoid(w.wave, dx=1.0, initial=0) * w.dt
    assert np.array_equal(ga, (area * 267.7 / 10000).astype(np.float32))
    assert np.array_equal(g, (w.wave * w.dt * 267.5 / 10000).astype(np.float32))


def test_walker_bvalue_is_discrete_q_sum():
    # b of the walker's phase = dt * sum (cumsum g)^2 ; for the unit waveform
    # it is the reference b-value with gamma 267.5 instead of 267.513
    w = wf.PGDiffWaveform(30.0, 10.0, 40.0, 0.02)
    g, _ = wf.phase_coefficients(w.wave, w.dt)
    b_walk = wf.bvalue_of_coefficients(g.astype(np.float64), w.dt)
    assert b_walk == pytest.approx(w.calculate_bvalue_from_wave() * (267.5 / 267.513) ** 2, rel=1e-6)


def test_pgse_shells_scale_b():
    p = PGSEShells(10.0, 40.0, 0.01, [1.0, 2.0, 3.0], fibonacci_hemisphere(64))
    assert p.nsteps == 5000 and p.nsignals == 192
    assert p.b_walk == pytest.approx(1.0, rel=1e-6)
    assert np.allclose(p.b_per_signal, np.repeat([1.0, 2.0, 3.0], 64), rtol=1e-6)


def test_direction_sets():
    f = fibonacci_hemisphere(64)
    assert np.allclose(np.linalg.norm(f, axis=0), 1) and (f[2] > 0).all()
    i6 = icosahedron_6()
    assert np.allclose(np.linalg.norm(i6, axis=0), 1)
    cos = np.abs(i6.T @ i6)[np.triu_indices(6, 1)]
    assert np.allclose(cos, 1 / np.sqrt(5))  # equiangular lines


def test_ogse_waveforms_are_balanced():
    w = wf.ApodizedCosineOGSEWaveform(N_cycles=2, T_duration=20.0, te=50.0, time_step=0.01)
    assert abs(w.wave.sum()) < 1e-6 * np.abs(w.wave).sum()
    assert np.abs(w.wave).max() == pytest.approx(1.0, abs=1e-3)
    info = w.get_waveform_info()
    assert info["frequency_ratio"] == pytest.approx(2.0)
    c = wf.CosineOGDiffWaveform(num_oscillations=1, frequency=50.0, gmax=1.0, te=60.0, time_step=0.01)
    assert np.abs(c.wave).max() <= 1.0 + 1e-12
