"""Alias of qdsim_cython.core.mesh_minimal (reference mesh_minimal.pyx)."""
from qdsim_amd.mesh import SimpleMesh, create_simple_mesh  # noqa: F401
