"""Command-line drivers of the walker: the narrow-pulse ADC/kurtosis
simulation and the DWI gradient simulation (cli/main.py of the reference)."""
