"""Drop-in counterpart of simulation_toolkit.simulation_engine: the walker's
host API (DiffSim3d, DwiSim3d), substrate model (SimGeometry3D, Structure3D)
and gradient waveforms, backed by libnozomi_hip.so."""
