"""Physical constants and unit conversions of SPILADY (spilady.h:53-62).

Internal units: length in Angstrom, energy in eV, time in s, mass in
eV/(Angstrom/s)^2, temperature as k_B*T in eV, spin dimensionless
(magnetic moment M = -g*muB*S).
"""

SPEED_OF_LIGHT = 299792458e0          # m/s
BOLTZ = 8.617343e-5                   # eV/K
UTOEV = 931.494028e6 / SPEED_OF_LIGHT / SPEED_OF_LIGHT   # eV/(m/s)^2 per u
LENGTH_SCALE = 1e-10                  # 1 Angstrom in m
HBAR = 6.58211899e-16                 # eV s
MUB = 5.7883817555e-5                 # eV/T
EL_G = 2.0023193043622e0              # electron g factor
PI_NUM = 3.141592653589793238462e0
EVTOJ = 1.602176487e-19

FE_MASS_U = 55.847
FE_A_LATTICE = 2.8665


def mass_internal(mass_u: float) -> float:
    """Atomic mass in eV/(A/s)^2 (read_variables.cpp:155-157)."""
    return mass_u * UTOEV * LENGTH_SCALE ** 2
