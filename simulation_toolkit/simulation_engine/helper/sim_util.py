from nozomi_amd.simulation_engine.helper.sim_util import validate_compartment_intra  # noqa: F401
