"""Alias of qdsim_cython.solvers.fixed_open_system_solver."""
from qdsim_amd.solvers import FixedOpenSystemSolver  # noqa: F401
