"""qdsim_amd -- the FEM-Hamiltonian path of QDSim, MI355X-native.

Python call surface mirroring qdsim_cython.solvers (FEMSolver, EigenSolver,
SelfConsistentSolver, FixedOpenSystemSolver, PoissonSolver) over hand-written
HIP kernels for gfx950 (libqdsim_hip.so, C-ABI in include/qdsim_hip.h).
"""
from . import _lib  # noqa: F401  (fails loudly when the native library is missing)
from .mesh import SimpleMesh, create_simple_mesh
from .mesh_io import FileMesh, load_csr, load_mesh, load_pattern, save_csr, save_mesh, save_pattern

__all__ = ["SimpleMesh", "create_simple_mesh", "FEMSolver", "EigenSolver", "SelfConsistentSolver",
           "FixedOpenSystemSolver", "PoissonSolver", "SchrodingerProblem", "create_fem_solver",
           "create_eigen_solver", "lifetimes_fs", "FEInterpolator", "create_interpolator",
           "FileMesh", "load_mesh", "save_mesh", "load_csr", "save_csr", "load_pattern", "save_pattern",
           "bias_sweep", "bias_sweep_instances", "gate_gradients", "gate_parameter_sets",
           "FullPoissonDriftDiffusionSolver", "Material"]


def __getattr__(name):
    # solver classes import torch; load them on first use
    if name in ("FEMSolver", "create_fem_solver"):
        from . import fem
        return getattr(fem, name)
    if name in ("EigenSolver", "create_eigen_solver"):
        from . import eigen
        return getattr(eigen, name)
    if name == "SelfConsistentSolver":
        from .self_consistent import SelfConsistentSolver
        return SelfConsistentSolver
    if name in ("FixedOpenSystemSolver", "lifetimes_fs"):
        from . import solvers
        return getattr(solvers, name)
    if name == "PoissonSolver":
        from .poisson import PoissonSolver
        return PoissonSolver
    if name in ("FEInterpolator", "create_interpolator"):
        from . import interpolator
        return getattr(interpolator, name)
    if name == "SchrodingerProblem":
        from .schrodinger import SchrodingerProblem
        return SchrodingerProblem
    if name in ("FullPoissonDriftDiffusionSolver", "Material"):
        from . import drift_diffusion
        return getattr(drift_diffusion, name)
    if name in ("bias_sweep", "bias_sweep_instances", "gate_gradients", "gate_parameter_sets"):
        from . import batch
        return getattr(batch, name)
    raise AttributeError(name)
