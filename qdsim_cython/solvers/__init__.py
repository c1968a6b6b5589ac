"""Aliases of qdsim_cython.solvers (reference solvers/__init__.py:51-65)."""


def __getattr__(name):
    import qdsim_amd
    aliases = {"CythonFEMSolver": "FEMSolver", "CythonEigenSolver": "EigenSolver",
               "CythonSelfConsistentSolver": "SelfConsistentSolver",
               "CythonPoissonSolver": "PoissonSolver"}
    return getattr(qdsim_amd, aliases.get(name, name))
