"""Writes waveform_fixtures.npz: the reference's sampled OGSE waveforms.

CosineOGDiffWaveform and ApodizedCosineOGSEWaveform
(simulation_toolkit/simulation_engine/waveforms.py:188-343) place their
samples with float comparisons against ramp and cosine boundaries; no closed
form or reference test pins those placements, so this script loads the
reference module from its file (numpy and scipy only; the package __init__,
which needs a CUDA device, is not imported) and records inputs and outputs.
Run in the development container only:

    python tests/golden/make_waveform_fixtures.py /root/reference

The PGSE waveform is not recorded: its sample count and lobe edges are pinned
by closed forms in tests/test_waveforms.py.
"""
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# (num_oscillations, frequency Hz, gmax, te ms, dt ms)
COSINE = [(1, 50.0, 1.0, 60.0, 0.01), (2, 100.0, 0.5, 50.0, 0.02), (3, 150.0, 1.0, 40.0, 0.005),
          (4, 25.0, 1.0, 400.0, 0.05), (2, 120.0, 0.8, 45.0, 0.003)]
# (N_cycles, T_duration ms, te ms or None, gmax, dt ms)
APODIZED = [(2, 10.0, None, 1.0, 0.01), (1, 20.0, 90.0, 0.5, 0.02), (3, 7.5, None, 1.0, 0.005),
            (4, 12.0, 60.0, 1.0, 0.01), (5, 9.0, 37.0, 0.7, 0.003)]


def load(ref_root):
    path = os.path.join(ref_root, "simulation_toolkit", "simulation_engine", "waveforms.py")
    spec = importlib.util.spec_from_file_location("reference_waveforms", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main(ref_root):
    wf = load(ref_root)
    out = {}
    for k, (n, f, gmax, te, dt) in enumerate(COSINE):
        w = wf.CosineOGDiffWaveform(n, f, gmax, te, time_step=dt)
        out[f"cosine_{k}_t_ends"] = np.array([w.t[0], w.t[-1], w.t.size], dtype=np.float64)
        out[f"cosine_{k}_wave"] = w.wave
    for k, (n, T, te, gmax, dt) in enumerate(APODIZED):
        w = wf.ApodizedCosineOGSEWaveform(n, T, te=te, gmax=gmax, time_step=dt)
        out[f"apodized_{k}_t_ends"] = np.array([w.t[0], w.t[-1], w.t.size], dtype=np.float64)
        out[f"apodized_{k}_wave"] = w.wave
        out[f"apodized_{k}_te"] = np.float64(w.te)
    out["params"] = np.frombuffer(json.dumps({"cosine": COSINE, "apodized": APODIZED}).encode(),
                                  dtype=np.uint8)
    dst = os.path.join(HERE, "waveform_fixtures.npz")
    np.savez_compressed(dst, **out)
    print(dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
