"""Physical constants and the internal unit system.

The public API is SI (metres, joules, kilograms), as in the reference Cython
solvers (qdsim_cython/solvers/fixed_open_system_solver.pyx:24-26,
schrodinger_solver.pyx:27-29; CODATA-2018 values).

Device kernels never see SI values: hbar^2 = 1.1e-68 J^2 s^2 underflows FP32 and
hbar^2/2m* ~ 9e-38 J m^2 sits at the FP32 normal limit.  Everything on the GPU is
expressed in nanometres and electron-volts, where hbar^2/(2 m0) = 0.0381 eV nm^2.
Conversion happens only at the Python boundary.
"""

HBAR = 1.054571817e-34          # J s
M_E = 9.1093837015e-31          # kg
Q_E = 1.602176634e-19           # C
EV_TO_J = 1.602176634e-19       # J / eV
K_B = 1.380649e-23              # J / K
EPSILON_0 = 8.854187817e-12     # F / m  (value used by poisson_solver.pyx:28)

NM = 1e-9                       # m per internal length unit

#: hbar^2/(2 m0) in eV nm^2
HB2M0_EV_NM2 = HBAR * HBAR / (2.0 * M_E) / EV_TO_J / (NM * NM)


def kinetic_coefficient_ev_nm2(m_star_kg):
    """hbar^2/(2 m*) in eV nm^2 for an effective mass given in kg (scalar or array)."""
    return (HBAR * HBAR / (2.0 * m_star_kg)) / EV_TO_J / (NM * NM)
