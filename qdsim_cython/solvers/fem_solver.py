"""Alias of qdsim_cython.solvers.fem_solver (reference fem_solver.pyx)."""
from qdsim_amd import FEMSolver as CythonFEMSolver  # noqa: F401
from qdsim_amd import FEMSolver  # noqa: F401
from qdsim_amd import create_fem_solver  # noqa: F401
