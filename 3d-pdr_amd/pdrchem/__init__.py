"""pdrchem: MI355X-native per-grid-point chemistry solve for 3D-PDR.

Host-side Python mirror of the reference's chemistry interface
(calculate_abundances.c:64, calc_reac_rates.F90:16) over the C-ABI library
``libpdrchem.so`` built from ``csrc/``.
"""
