"""Alias of qdsim_cython.solvers.eigen_solver (reference eigen_solver.pyx)."""
from qdsim_amd import EigenSolver as CythonEigenSolver  # noqa: F401
from qdsim_amd import EigenSolver  # noqa: F401
from qdsim_amd import create_eigen_solver  # noqa: F401
