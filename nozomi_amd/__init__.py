"""nozomi_amd: MI355X (gfx950) Monte Carlo diffusion-MRI random walker.

The walk, seeding, phase and signal kernels live in libnozomi_hip.so (HIP,
C ABI in include/nozomi_hip.h); this package is the Python host side that
mirrors the reference's simulation_engine API.  The drop-in import paths of
the reference (simulation_toolkit.*) are provided by the sibling package
simulation_toolkit/.
"""
__version__ = "1.0.0"
