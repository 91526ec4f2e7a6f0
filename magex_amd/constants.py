"""Physical constants in magex's units (magex/spinmodel.py:28-30): meV, T, ps, K."""
MU_B = 5.7883818060e-2      # Bohr magneton, meV / T
K_B = 8.617333262e-2        # Boltzmann constant, meV / K
GAMMA = 0.1760859630        # electron gyromagnetic ratio, rad / (ps T)
