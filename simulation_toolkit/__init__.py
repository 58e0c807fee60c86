"""Import-path compatibility package: the reference's module paths
(simulation_toolkit.simulation_engine.diffsim3d, ...) resolved to nozomi_amd.

Unlike the reference (simulation_toolkit/__init__.py:12), importing the
package does not touch the GPU."""
from nozomi_amd import __version__  # noqa: F401
