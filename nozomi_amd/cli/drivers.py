"""Simulation drivers over a substrate file: the narrow-pulse ADC/kurtosis
driver (cli/simulation_main.py:19-149 of the reference) and the DWI driver the
reference's CLI imports but does not ship (cli/main.py:17,72 name
``simulation_gradient_main``; its protocol is experiment/setup/simulation/
pgse_sim_params.json and ogse_sim_params.json).

Differences from the reference, each a fix of a reference defect:
- every time step's displacement moments are accumulated in FP64 inside the
  walk launch instead of one float-atomic kernel per step
  (sim3d_kernel.cu:162-238);
- the ADC output pairs D(t_k) with t_k = k*dt, k = 1..nt; the reference pairs
  D[1:] with diff_time[1:], one step late, and its last time is 0
  (simulation_main.py:101,110);
- the walk is seeded (``seed`` in the parameters, else drawn), while the
  reference's seeding is unseeded (geometry.py:148).
"""
import os
import time

import numpy as np
import torch

from .. import substrate as sb
from ..protocol import OGSEShells, PGSEShells, fibonacci_hemisphere
from ..simulation_engine import diffsim3d as ds3
from ..simulation_engine import geometry as geom


def substrate_geometry(substrate_file, D, T2=100.0, rho=1.0):
    """Reference substrate pickle -> SimGeometry3D with one structure per
    z-padded fiber (simulation_main.py:35-54).  Returns (sg3, nfibers, L)."""
    fibers_matrix, L = sb.import_array_geometry_full_path(substrate_file)
    fibers = sb.createBoundaryZ(sb.split_matrix_to_list(fibers_matrix, L), L)
    sg3 = geom.SimGeometry3D(L, L, L, D, T2, rho)
    for f in fibers:
        sg3.add_structure(geom.Structure3D(f[:, 0], f[:, 1], f[:, 2], f[:, 3], D, T2, rho))
    return sg3, len(fibers), L


def _compartment_D(params, compartment):
    if compartment == "intra":
        return float(params["D0_intra"])
    if compartment == "extra":
        return float(params["D0_extra"])
    raise ValueError(f"compartment must be 'intra' or 'extra', not {compartment!r}")


def _targets(compartment, nfibers):
    """Seeding targets (simulation_main.py:63-66): every fiber, or the box."""
    return list(range(nfibers)) if compartment == "intra" else [nfibers]


def _seed(params):
    s = params.get("seed")
    return int(np.random.randint(np.iinfo(np.int32).max)) if s is None else int(s)


def _out_folder(substrate_file, *sub):
    root = os.path.dirname(os.path.dirname(os.path.abspath(substrate_file)))
    path = os.path.join(root, "sim", *sub)
    os.makedirs(path, exist_ok=True)
    return path


def _save_pickle(path, data):
    import pickle
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        pickle.dump(data, f, protocol=pickle.HIGHEST_PROTOCOL)


def plot_ADC_vs_time(diff_time, Dx, Dy, Dz, folder_name, file_name):
    """PNG of D_x, D_y, D_xy, D_z against diffusion time (semilog), as
    simulation_report.py:15-35 draws it.  Returns the file path."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, ax = plt.subplots()
    ax.semilogx(diff_time, Dx, "b.", markersize=10)
    ax.semilogx(diff_time, Dy, "y.", markersize=10)
    ax.semilogx(diff_time, (Dx + Dy) / 2, "g.", markersize=10)
    ax.semilogx(diff_time, Dz, "r.", markersize=10)
    ax.set_xlabel("t (ms)")
    ax.set_ylabel("D (µm$^2$/ms)")
    ax.set_title(" Diffusion coefficients vs. diffusion time ", fontsize=17)
    ax.grid(True)
    ax.set_xlim([0.0001, max(diff_time)])
    ax.set_ylim([0.0, 3.2])
    ax.legend(["Dx", "Dy", "Dxy", "Dz"], fontsize=18)
    os.makedirs(folder_name, exist_ok=True)
    path = os.path.join(folder_name,
                        "difftime_" + str(np.round(np.max(diff_time), 2)) + "_" + file_name + ".png")
    fig.savefig(path)
    plt.close(fig)
    return path


def adc_from_step_moments(stepmom, nspins, dt):
    """Per-step D_a = <da^2>/(2t) and K_a = <da^4>/<da^2>^2 - 3 from the FP64
    step moment sums [nt, 6] (sim3d_kernel.cu:162-238, simulation_main.py:
    104-112); t_k = k*dt.  Returns (t [nt], D [nt, 3], K [nt, 3])."""
    s = np.asarray(stepmom, dtype=np.float64)
    t = dt * np.arange(1, s.shape[0] + 1)
    m2 = s[:, :3] / nspins
    m4 = s[:, 3:] / nspins
    D = m2 / (2.0 * t[:, None])
    with np.errstate(divide="ignore", invalid="ignore"):
        K = m4 / (m2 * m2) - 3.0
    return t, D, K


def simulation_main(params, substrate_file, device=None, plot=True, chunk=1000, log=print):
    """Narrow-pulse driver: walk sim_time/time_step steps with the T2 weight
    and write the ADC pickle [nt, 4] = (Dx, Dy, Dz, t) and its plot under
    <substrate root>/sim (simulation_main.py:19-149).  Returns a dict with
    the arrays and output paths."""
    total_sim_time = float(params["sim_time"])
    time_step = float(params["time_step"])
    num_spins = int(params["num_spins"])
    compartment = params["compartment"]
    nseg = int(params["nseg"])
    D = _compartment_D(params, compartment)
    sg3, nfibers, L = substrate_geometry(substrate_file, D)
    log(f"box length {L}")
    nt = int(total_sim_time / time_step)
    seed = _seed(params)

    t0 = time.time()
    sim = ds3.DiffSim3d(sg3, num_spins, device=device)
    sim.set_segments(nsegx=nseg, nsegy=nseg, nsegz=nseg)
    sim.setup(structures=_targets(compartment, nfibers), initstates_int=seed)
    start = sim.spins
    torch.cuda.synchronize(sim.device)
    table_elapsed = np.round(time.time() - t0, 2)

    t1 = time.time()
    stepmom = torch.zeros((nt, 6), dtype=torch.float64, device=sim.device)
    done = 0
    while done < nt:
        k = min(chunk, nt - done)
        sim.run(k, time_step, stepmom=stepmom[done:done + k])
        done += k
        if done % max(chunk, nt // 10 or 1) < k or done == nt:
            log(f"Progress: {100.0 * done / nt:.1f}% | Time: {done * time_step:.3f} ms")
    mom = stepmom.cpu().numpy()
    elapsed = np.round(time.time() - t1, 2)
    t, Dk, Kk = adc_from_step_moments(mom, num_spins, time_step)

    base = os.path.splitext(os.path.basename(substrate_file))[0]
    name = (f"diffcoeff_{compartment}_{nseg}segments_{num_spins}spins_{base}_"
            f"TABLEtime{table_elapsed}sec_SIMtime{round(float(elapsed), 2)}sec_timestep{time_step}ms")
    folder = _out_folder(substrate_file)
    data_path = os.path.join(folder, "ADCdata", name + ".pkl")
    _save_pickle(data_path, np.column_stack((Dk[:, 0], Dk[:, 1], Dk[:, 2], t)))
    kurt_path = os.path.join(folder, "ADCdata", "kurtosis_" + name + ".pkl")
    _save_pickle(kurt_path, np.column_stack((Kk[:, 0], Kk[:, 1], Kk[:, 2], t)))
    png = plot_ADC_vs_time(t, Dk[:, 0], Dk[:, 1], Dk[:, 2], folder, name) if plot else None
    log("Simulation completed!")
    return {"t": t, "D": Dk, "K": Kk, "stepmom": mom, "seed": seed, "start": start,
            "data_path": data_path, "kurtosis_path": kurt_path, "png": png,
            "sim_seconds": float(elapsed)}


def gradient_protocol(params):
    """The DWI protocol of a gradient simulation config: the waveform of
    ``gradient.type`` (PGSE: big_delta, little_delta; OGSE: num_oscil cycles
    in a lobe of ``duration``) over te = sim_time on the time_step grid, at
    ``gradient.b_values`` (ms/um^2, default 1.0) along ``gradient.direction``
    (one [3] vector or a list of them) or ``gradient.num_directions``
    Fibonacci-hemisphere directions."""
    g = params["gradient"]
    dt = float(params["time_step"])
    te = float(params["sim_time"])
    b = g.get("b_values", [1.0])
    if g.get("num_directions"):
        dirs = fibonacci_hemisphere(int(g["num_directions"]))
    else:
        d = np.asarray(g.get("direction", [1, 0, 0]), dtype=np.float64)
        dirs = d.reshape(3, 1) if d.ndim == 1 else d.T
    kind = str(g["type"]).upper()
    if kind == "PGSE":
        return PGSEShells(g["little_delta"], g["big_delta"], dt, b, dirs, te=te)
    if kind == "OGSE":
        return OGSEShells(g["num_oscil"], g["duration"], te, dt, b, dirs)
    raise ValueError(f"gradient type must be PGSE or OGSE, not {g['type']!r}")


def write_fsl_bvals_bvecs(prefix, bvals_s_mm2, bvecs):
    """FSL text files: prefix.bval (one row, s/mm^2) and prefix.bvec (3 rows)."""
    np.savetxt(prefix + ".bval", np.asarray(bvals_s_mm2)[None, :], fmt="%.6g")
    np.savetxt(prefix + ".bvec", np.asarray(bvecs), fmt="%.8f")
    return prefix + ".bval", prefix + ".bvec"


def simulation_gradient_main(params, substrate_file, device=None, log=print):
    """DWI driver: walk the whole waveform with the phase moment, form every
    (b-shell, direction) signal in FP64 on the device and write the signals
    with their b-values and directions under <substrate root>/sim/DWIdata.
    Returns a dict with the signals and output paths."""
    num_spins = int(params["num_spins"])
    compartment = params["compartment"]
    nseg = int(params["nseg"])
    D = _compartment_D(params, compartment)
    sg3, nfibers, L = substrate_geometry(substrate_file, D)
    prot = gradient_protocol(params)
    seed = _seed(params)

    sim = ds3.DwiSim3d(sg3, num_spins, diffdir=prot.dirs, device=device)
    sim.set_segments(nsegx=nseg, nsegy=nseg, nsegz=nseg)
    t0 = time.time()
    sim.setup(structures=_targets(compartment, nfibers), initstates_int=seed)
    start = sim.spins
    g = torch.from_numpy(prot.g).to(sim.device)
    ga = torch.from_numpy(prot.garea).to(sim.device)
    sim._walk(prot.nsteps, prot.dt, g=g, garea=ga, t2=False)
    signals = sim.signals(prot.scales)
    elapsed = round(time.time() - t0, 2)

    kind = str(params["gradient"]["type"]).upper()
    base = os.path.splitext(os.path.basename(substrate_file))[0]
    name = f"dwi_{kind}_{compartment}_{nseg}segments_{num_spins}spins_{base}_timestep{prot.dt}ms"
    folder = _out_folder(substrate_file, "DWIdata")
    bvals = prot.b_per_signal * 1000.0
    data = {"signals": signals, "bvals_s_mm2": bvals, "bvecs": prot.dirs, "b_values": prot.b_values,
            "b_walk": prot.b_walk, "directions": prot.unit_dirs, "gradient": params["gradient"],
            "time_step": prot.dt, "nsteps": prot.nsteps, "num_spins": num_spins, "seed": seed,
            "box_length": L, "sim_seconds": elapsed}
    data_path = os.path.join(folder, name + ".pkl")
    _save_pickle(data_path, data)
    np.savetxt(os.path.join(folder, name + ".signal"), np.abs(signals)[None, :], fmt="%.8f")
    write_fsl_bvals_bvecs(os.path.join(folder, name), bvals, prot.dirs)
    log(f"DWI simulation completed: {prot.nsignals} signals, {prot.nsteps} steps, {elapsed} s")
    return {"signals": signals, "protocol": prot, "seed": seed, "start": start,
            "data_path": data_path}
