"""Alias of qdsim_cython.solvers.self_consistent_solver."""
from qdsim_amd.self_consistent import SelfConsistentSolver  # noqa: F401
from qdsim_amd.self_consistent import SelfConsistentSolver as CythonSelfConsistentSolver  # noqa: F401
