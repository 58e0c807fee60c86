"""validate_compartment_intra (simulation_engine/helper/sim_util.py:3-7)."""
import numpy as np


def validate_compartment_intra(sim, sg3):
    """True when every spin of sim lies inside one of sg3's structures (the
    box excluded)."""
    pos = sim.spins_d.get()
    inside = np.zeros(sim.nspins, dtype=bool)
    for structure in sg3.structures[:-1]:
        inside = np.logical_or(inside, structure.isinside(pos))
    return bool(inside.all())
