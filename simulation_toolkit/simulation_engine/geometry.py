from nozomi_amd.simulation_engine.geometry import SimGeometry3D, Structure3D  # noqa: F401
