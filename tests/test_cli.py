"""CLI and driver host logic (cli/main.py, cli/simulation_main.py of the
reference, and the DWI driver its CLI names): config merging, substrate
discovery, protocol construction and the ADC/kurtosis formulas."""
import json
import os

import numpy as np
import pytest

from nozomi_amd.cli import drivers, main as cli
from nozomi_amd.simulation_engine import waveforms as wf


def _write_substrate(tmp_path, axons=4, box=20.0, seed=5):
    rc = cli.main(["synthetic_substrate", "--output_folder_path", str(tmp_path), "--name", "exp",
                   "--axons", str(axons), "--box", str(box), "--seed", str(seed)])
    assert rc == 0
    files = cli.substrate_files(str(tmp_path))
    assert len(files) == 1
    return files[0]


def test_synthetic_substrate_round_trip(tmp_path):
    path = _write_substrate(tmp_path)
    sg3, nfib, L = drivers.substrate_geometry(path, D=2.0)
    assert L == 20.0 and nfib == sg3.nstructures >= 4
    # every fiber carries 25 z-images per end (adjust_geometry.py:22-36)
    sids = sg3.spheres[4]
    for s in range(nfib):
        z = sg3.spheres[2][sids == s]
        assert z.min() < -L / 2 and z.max() > L / 2
    assert sg3.structure_D().tolist() == [2.0] * (nfib + 1)
    assert sg3.structure_T2()[0] == 100.0


def test_config_merge_keeps_unset_values():
    args = cli.build_parser().parse_args(["simulation", "--substrates", "x", "--sim_time", "5"])
    conf = cli._merge(cli.load_config_file(args.config), args,
                      ("substrates", "sim_time", "compartment", "nseg", "num_spins", "time_step"))
    assert conf["substrates"] == "x" and conf["sim_time"] == 5.0
    assert conf["nseg"] == 20 and conf["time_step"] == 0.002 and conf["num_spins"] == 500000
    assert "compartment" not in conf


def test_substrate_listing_warns_without_data(tmp_path, capsys):
    (tmp_path / "a" / "data").mkdir(parents=True)
    (tmp_path / "a" / "data" / "s1.pkl").write_bytes(b"")
    (tmp_path / "b").mkdir()
    files = cli.substrate_files(str(tmp_path))
    assert [os.path.basename(f) for f in files] == ["s1.pkl"]
    assert "Data folder not found" in capsys.readouterr().out


def test_run_simulation_rejects_missing_folder(tmp_path):
    assert cli.run_simulation({"substrates": str(tmp_path / "nope")}) is None


def test_gradient_protocol_pgse_and_ogse():
    conf = cli.load_config_file(cli.GRADIENT_CONFIG_FILE)
    conf["gradient"]["b_values"] = [1.0, 2.5]
    p = drivers.gradient_protocol(conf)
    ref = wf.PGDiffWaveform(40.0, 10.0, 78.0, 0.002)
    assert p.nsteps == ref.wave.size
    assert np.array_equal(np.sign(p.wave), np.sign(ref.wave))
    assert p.nsignals == 2 and np.allclose(p.b_per_signal, [1.0, 2.5], rtol=1e-6)
    assert np.allclose(p.dirs, [[1, 1], [0, 0], [0, 0]])
    conf = json.load(open(os.path.join(cli.HERE, "setup", "ogse_sim_params.json")))
    conf["gradient"]["num_directions"] = 6
    p = drivers.gradient_protocol(conf)
    assert p.nsignals == 6 and p.b_walk == pytest.approx(1.0, rel=1e-6)
    assert abs(p.wave.sum()) < 1e-6 * np.abs(p.wave).sum()
    with pytest.raises(ValueError):
        drivers.gradient_protocol({**conf, "gradient": {"type": "STEAM"}})


def test_adc_from_step_moments_formulas():
    # Gaussian displacements with variance 2 D t per axis: D back, K = 0
    rng = np.random.default_rng(0)
    N, T, dt, D = 200000, 4, 0.5, np.array([1.0, 2.0, 0.5])
    sm = np.zeros((T, 6))
    for k in range(T):
        t = (k + 1) * dt
        d = rng.normal(size=(3, N)) * np.sqrt(2 * D * t)[:, None]
        sm[k, :3] = (d ** 2).sum(1)
        sm[k, 3:] = (d ** 4).sum(1)
    t, Dk, Kk = drivers.adc_from_step_moments(sm, N, dt)
    assert np.allclose(t, dt * np.arange(1, T + 1))
    assert np.allclose(Dk, D[None, :], rtol=2e-2)
    assert np.all(np.abs(Kk) < 0.05)


def test_fsl_files(tmp_path):
    b, v = drivers.write_fsl_bvals_bvecs(str(tmp_path / "x"), np.array([1000.0, 2000.0]),
                                         np.eye(3)[:, :2])
    assert np.allclose(np.loadtxt(b), [1000, 2000])
    assert np.loadtxt(v).shape == (3, 2)
