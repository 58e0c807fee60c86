from nozomi_amd.simulation_engine.waveforms import *  # noqa: F401,F403
from nozomi_amd.simulation_engine.waveforms import (  # noqa: F401
    ApodizedCosineOGSEWaveform, CosineOGDiffWaveform, DiffGradWaveform, PGDiffWaveform, dt0, gamma)
