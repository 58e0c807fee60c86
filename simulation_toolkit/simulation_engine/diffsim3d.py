from nozomi_amd.simulation_engine.diffsim3d import DiffSim3d, DwiSim3d  # noqa: F401
