"""Command-line entry point with the reference's simulation subcommands
(cli/main.py:100-179):

  python -m nozomi_amd.cli.main simulation --substrates DIR --sim_time T \
      --compartment intra|extra [--nseg N] [--config default-sim.json]
  python -m nozomi_amd.cli.main gradient_simulation --substrates DIR \
      [--sim_time T] [--compartment ...] [--nseg N] [--config pgse_sim_params.json]

DIR holds one folder per experiment, each with a data/ folder of substrate
pickles (cli/main.py:34-75); results go to <experiment>/sim/.  Command-line
values override the JSON config; unlike the reference (cli/main.py:146-149,
161-165), options left unset keep the config's values instead of replacing
them with None.  Substrate generation (the reference's ``substrate``
subcommand) is not part of this package; ``synthetic_substrate`` writes a
seeded beaded-axon substrate pickle in the reference's format instead.
"""
import argparse
import datetime
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SIM_CONFIG_FILE = os.path.join(HERE, "setup", "default-sim.json")
GRADIENT_CONFIG_FILE = os.path.join(HERE, "setup", "pgse_sim_params.json")


def load_config_file(path):
    """JSON config (config.py:13-18)."""
    with open(path) as f:
        return json.load(f)


def substrate_files(main_folder):
    """Every file under <main_folder>/<experiment>/data/, in listing order
    (cli/main.py:41-75)."""
    out = []
    for sub in sorted(os.listdir(main_folder)):
        data = os.path.join(main_folder, sub, "data")
        if os.path.isdir(os.path.join(main_folder, sub)) and os.path.isdir(data):
            out += [os.path.join(data, f) for f in sorted(os.listdir(data))
                    if os.path.isfile(os.path.join(data, f))]
        elif os.path.isdir(os.path.join(main_folder, sub)):
            print(f"Warning: Data folder not found in {os.path.join(main_folder, sub)}")
    return out


def run_simulation(sim_config, pulse="narrow", device=None):
    """Run the narrow-pulse or gradient driver on every substrate
    (cli/main.py:34-75).  Returns the list of driver results."""
    from . import drivers
    main_folder = sim_config["substrates"]
    if not main_folder or not os.path.isdir(main_folder):
        print(f"Error: {main_folder} is not a valid directory")
        return None
    files = substrate_files(main_folder)
    results = []
    for k, path in enumerate(files, 1):
        print(f"=====Starting simulation {k}/{len(files)}=====")
        if pulse == "narrow":
            results.append(drivers.simulation_main(sim_config, path, device=device))
        else:
            results.append(drivers.simulation_gradient_main(sim_config, path, device=device))
    print(f"Simulation completed for {os.path.basename(os.path.normpath(main_folder))}")
    return results


def _merge(config, args, keys):
    for k in keys:
        v = getattr(args, k, None)
        if v is not None:
            config[k] = v
    return config


def _save_config(sim_config, config_path):
    stamp = datetime.datetime.now().strftime("%Y%m%d_%H%M%S")
    out = os.path.join(sim_config["substrates"], stamp + "_" + os.path.basename(config_path))
    with open(out, "w") as f:
        json.dump(sim_config, f, indent=4)
    return out


def build_parser():
    p = argparse.ArgumentParser(description="Diffusion MRI Monte Carlo walker (MI355X)")
    sp = p.add_subparsers(dest="command", help="Available commands")
    for name, default in (("simulation", SIM_CONFIG_FILE),
                          ("gradient_simulation", GRADIENT_CONFIG_FILE)):
        s = sp.add_parser(name, help="Run simulations" if name == "simulation"
                          else "Run gradient simulations")
        s.add_argument("--substrates", type=str, help="Path to substrates folder")
        s.add_argument("--sim_time", type=float, help="Total diffusion time (ms)")
        s.add_argument("--compartment", type=str, help="Compartment to simulate (intra, extra)")
        s.add_argument("--nseg", type=int, help="Number of segments per axis (cell-size hint)")
        s.add_argument("--num_spins", type=int, help="Number of spins")
        s.add_argument("--time_step", type=float, help="Time step (ms)")
        s.add_argument("--seed", type=int, help="Philox key of the walk")
        s.add_argument("--config", type=str, default=default, help="JSON configuration file")
        s.add_argument("--device", type=int, default=None, help="GPU index")
    s = sp.add_parser("synthetic_substrate", help="Write a seeded beaded-axon substrate pickle")
    s.add_argument("--output_folder_path", type=str, required=True)
    s.add_argument("--name", type=str, default="beaded")
    s.add_argument("--axons", type=int, default=16)
    s.add_argument("--box", type=float, default=30.0)
    s.add_argument("--seed", type=int, default=20261203)
    return p


def main(argv=None):
    args = build_parser().parse_args(argv)
    if not args.command:
        build_parser().print_help()
        return 0
    if args.command == "synthetic_substrate":
        from ..substrate import generate_beaded_axons, save_data_array_to_pickle
        M, L = generate_beaded_axons(args.axons, args.box, seed=args.seed)
        path = os.path.join(args.output_folder_path, args.name, "data",
                            f"{args.name}_{args.axons}axons_L{args.box:g}_seed{args.seed}.pkl")
        os.makedirs(os.path.dirname(path), exist_ok=True)
        save_data_array_to_pickle(path, M, L)
        print(path)
        return 0
    keys = ("substrates", "sim_time", "compartment", "nseg", "num_spins", "time_step", "seed")
    sim_config = _merge(load_config_file(args.config), args, keys)
    import torch
    device = None if args.device is None else torch.device("cuda", args.device)
    res = run_simulation(sim_config, "narrow" if args.command == "simulation" else "wide", device)
    if res is None:
        return 1
    _save_config(sim_config, args.config)
    return 0


if __name__ == "__main__":
    sys.exit(main())
