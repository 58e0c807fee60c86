"""The narrow-pulse ADC driver and the DWI gradient driver end to end on a
synthetic substrate pickle, against the FP64 oracle walked from the same
start positions with the same Philox key (trajectories are the oracle's, so
the bars are round-off bars)."""
import os
import pickle

import numpy as np
import pytest

from nozomi_amd.cli import drivers, main as cli

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def substrate_file(tmp_path_factory):
    root = tmp_path_factory.mktemp("subs")
    assert cli.main(["synthetic_substrate", "--output_folder_path", str(root), "--name", "exp",
                     "--axons", "16", "--box", "30", "--seed", "20261203"]) == 0
    return cli.substrate_files(str(root))[0]


def _oracle(sg3, oracle_lib):
    return oracle_lib.OracleSubstrate(sg3.spheres, sg3.nstructures, sg3.structure_D(),
                                      sg3.structure_T2(), (sg3.Lx, sg3.Ly, sg3.Lz))


@pytest.mark.parametrize("compartment", ["intra", "extra"])
def test_narrow_pulse_driver_matches_oracle(substrate_file, oracle_lib, compartment):
    params = {"sim_time": 4.0, "time_step": 0.02, "num_spins": 8192, "compartment": compartment,
              "D0_intra": 2.25, "D0_extra": 2.0, "nseg": 20, "seed": 4242}
    print(f"narrow-pulse driver {compartment}", flush=True)
    res = drivers.simulation_main(params, substrate_file, chunk=64, log=lambda *a: None)
    assert os.path.isfile(res["data_path"]) and os.path.isfile(res["png"])
    with open(res["data_path"], "rb") as f:
        table = pickle.load(f)
    assert table.shape == (200, 4)
    assert np.allclose(table[:, 3], 0.02 * np.arange(1, 201))
    sg3, _, _ = drivers.substrate_geometry(substrate_file, params[f"D0_{compartment}"])
    orc = _oracle(sg3, oracle_lib)
    p = res["start"].astype(np.float64)
    out = orc.walk(p.copy(), p.copy(), np.ones(p.shape[1]), seed=res["seed"], step0=0, nsteps=200,
                   dt=0.02, stepmom=True)
    _, D_ref, _ = drivers.adc_from_step_moments(out["stepmom"], p.shape[1], 0.02)
    rel = np.abs(table[:, :3] / D_ref - 1)
    print(f"ADC vs oracle: max rel {rel.max():.2e}")
    # equal trajectories and FP64 moment sums: only the summation order and
    # the table's storage precision separate the two
    assert rel.max() < 1e-6
    if compartment == "intra":
        # restricted: the transverse ADC falls well below D0 within 4 ms
        assert table[-1, :2].mean() < 0.6 * params["D0_intra"]


def test_gradient_driver_matches_oracle(substrate_file, oracle_lib):
    params = {"sim_time": 20.0, "time_step": 0.02, "num_spins": 32768, "compartment": "extra",
              "D0_intra": 2.0, "D0_extra": 2.0, "nseg": 20, "seed": 77,
              "gradient": {"type": "PGSE", "big_delta": 10.0, "little_delta": 5.0,
                           "b_values": [1.0, 2.0], "num_directions": 6}}
    print("gradient driver PGSE", flush=True)
    res = drivers.simulation_gradient_main(params, substrate_file, log=lambda *a: None)
    prot = res["protocol"]
    base = res["data_path"][:-4]
    assert np.allclose(np.loadtxt(base + ".bval"), prot.b_per_signal * 1000)
    sg3, _, _ = drivers.substrate_geometry(substrate_file, 2.0)
    orc = _oracle(sg3, oracle_lib)
    p = res["start"].astype(np.float64)
    N = p.shape[1]
    mom = np.zeros((3, N))
    orc.walk(p.copy(), p.copy(), np.ones(N), seed=res["seed"], step0=0, nsteps=prot.nsteps,
             dt=prot.dt, g=prot.g.astype(np.float64), garea=prot.garea.astype(np.float64), mom=mom)
    ph = (prot.dirs.T @ mom) * prot.scales[:, None]
    ref = np.exp(1j * ph).mean(1)
    err = np.abs(res["signals"] - ref)
    print(f"signals vs oracle: max abs {err.max():.2e}")
    assert err.max() < 1e-3
    assert err.max() < 1e-6  # equal trajectories: FP32 moments are all that differ
    # extra-axonal signal decays with b and stays above free diffusion's
    mag = np.abs(res["signals"]).reshape(2, 6)
    assert np.all(mag[1] < mag[0]) and np.all(mag[0] > np.exp(-1.0 * 2.0) - 0.01)
