"""Alias of qdsim_cython.solvers.poisson_solver."""
from qdsim_amd.poisson import PoissonSolver  # noqa: F401
from qdsim_amd.poisson import PoissonSolver as CythonPoissonSolver  # noqa: F401
