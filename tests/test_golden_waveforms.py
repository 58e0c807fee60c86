"""Waveform sample placement against the reference's own outputs: the
vectors in tests/golden/waveform_fixtures.npz were written by
tests/golden/make_waveform_fixtures.py from the reference's
CosineOGDiffWaveform and ApodizedCosineOGSEWaveform
(waveforms.py:188-343).  Bit-for-bit: the number of samples, their times and
values decide the walk's steps and phases."""
import json
import os

import numpy as np
import pytest

from nozomi_amd.simulation_engine import waveforms as wf

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "waveform_fixtures.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(FIX, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    data["params"] = json.loads(bytes(data["params"]).decode())
    return data


def _ends(w):
    return np.array([w.t[0], w.t[-1], w.t.size], dtype=np.float64)


def test_cosine_ogse_matches_reference(golden):
    for k, (n, f, gmax, te, dt) in enumerate(golden["params"]["cosine"]):
        w = wf.CosineOGDiffWaveform(n, f, gmax, te, time_step=dt)
        assert np.array_equal(_ends(w), golden[f"cosine_{k}_t_ends"]), k
        assert np.array_equal(w.wave, golden[f"cosine_{k}_wave"]), k
        assert np.count_nonzero(w.wave) > 10


def test_apodized_ogse_matches_reference(golden):
    for k, (n, T, te, gmax, dt) in enumerate(golden["params"]["apodized"]):
        w = wf.ApodizedCosineOGSEWaveform(n, T, te=te, gmax=gmax, time_step=dt)
        assert w.te == golden[f"apodized_{k}_te"], k
        assert np.array_equal(_ends(w), golden[f"apodized_{k}_t_ends"]), k
        assert np.array_equal(w.wave, golden[f"apodized_{k}_wave"]), k
        assert np.count_nonzero(w.wave) > 10
