"""Unit conversions used by the MD surface (deepmd_jax/md.py:24-26, deepmd_jax/pimd.py:19)."""
MASS_UNIT_CONVERSION = 1.036427e2   # Dalton -> eV fs^2 / Angstrom^2
TEMP_UNIT_CONVERSION = 8.617333e-5  # Kelvin -> eV
PRESS_UNIT_CONVERSION = 6.241509e-7  # bar -> eV / Angstrom^3
HBAR = 6.582119569e-16 * 1e15       # eV fs
