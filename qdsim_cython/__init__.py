"""Import-path compatibility for code written against qdsim_cython
(qdsim_cython/__init__.py:24-47, solvers/__init__.py:51-65 of the reference).
Everything resolves to the qdsim_amd GPU implementation."""
from qdsim_amd import SimpleMesh  # noqa: F401


def __getattr__(name):
    import qdsim_amd
    aliases = {"CythonFEMSolver": "FEMSolver", "CythonEigenSolver": "EigenSolver",
               "CythonSelfConsistentSolver": "SelfConsistentSolver",
               "CythonPoissonSolver": "PoissonSolver"}
    return getattr(qdsim_amd, aliases.get(name, name))
