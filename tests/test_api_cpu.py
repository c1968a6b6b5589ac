"""Call-surface behaviour that needs no GPU: constructors, attributes, presets,
argument errors (fixed_open_system_solver.pyx:28-376, eigen_solver.pyx:40-68,
mesh_minimal.pyx)."""
import numpy as np
import pytest

EV = 1.602176634e-19
M_E = 9.1093837015e-31


def test_fixed_open_system_constructor_and_presets():
    from qdsim_amd.solvers import FixedOpenSystemSolver, lifetimes_fs
    s = FixedOpenSystemSolver(8, 6, 25e-9, 20e-9, lambda x, y: 0.067 * M_E, lambda x, y: 0.0, True)
    assert (s.nx, s.ny, s.num_nodes, s.num_elements) == (8, 6, 48, 70)
    assert s.use_open_boundaries and s.dirac_normalization
    assert s.cap_strength == pytest.approx(0.01 * EV) and s.cap_length_ratio == 0.2
    assert s.elements.shape == (70, 3) and len(s.nodes_x) == 48
    s.configure_device_specific_solver("quantum_well")
    assert s.cap_strength == pytest.approx(0.02 * EV) and s.cap_length_ratio == 0.25
    s.configure_device_specific_solver("quantum_dot", {"cap_strength": 1e-22})
    assert s.cap_strength == 1e-22 and s.cap_length_ratio == 0.1
    s.apply_conservative_boundary_conditions()
    assert s.cap_strength == pytest.approx(0.001 * EV) and s.cap_length_ratio == 0.05
    s.apply_minimal_cap_boundaries()
    assert s.cap_strength == pytest.approx(0.002 * EV)
    assert s.get_eigenvalues_eV().size == 0
    # CAP profile of fixed_open_system_solver.pyx:124-142
    cl = 0.1 * 20e-9
    assert s.cap_absorption(np.array([0.0]), np.array([0.0]))[0] == pytest.approx(s.cap_strength)
    assert s.cap_absorption(np.array([cl]), np.array([0.0]))[0] == pytest.approx(0.0)
    assert s.cap_absorption(np.array([12e-9]), np.array([0.0]))[0] == 0.0
    tau = lifetimes_fs(np.array([1e-21 - 1j * 1e-22]))
    assert tau[0] == pytest.approx(1.054571817e-34 / 2e-22 * 1e15)


def test_eigensolver_defaults_and_analytic():
    from qdsim_amd.eigen import EigenSolver
    from qdsim_amd.mesh import SimpleMesh
    es = EigenSolver(SimpleMesh(5, 4))
    assert es.matrix_size == 20
    assert es.solver_options == {'tolerance': 1e-8, 'max_iterations': 1000, 'which': 'SM',
                                 'sigma': None, 'mode': 'normal'}
    es.set_solver_options(tolerance=1e-6)
    assert es.solver_options['tolerance'] == 1e-6
    hbar = 1.054571817e-34
    assert np.allclose(es.solve_quantum_harmonic_oscillator(3, 1e12, M_E), hbar * 1e12 * np.array([0.5, 1.5, 2.5]))
    box = es.solve_particle_in_box(3, 10e-9, M_E)
    assert np.allclose(box / box[0], [1, 4, 9])
    assert es.get_last_solution() == (None, None)


def test_mesh_errors():
    from qdsim_amd.mesh import SimpleMesh
    with pytest.raises(ValueError):
        SimpleMesh(5, 4, 1e-8, 1e-8, 4)
    m = SimpleMesh(5, 4)
    with pytest.raises(IndexError):
        m.get_node_at(20)


def test_fem_solver_rejects_unknown_method_without_gpu_work():
    # the method check happens before any device call (fem_solver.pyx:391)
    from qdsim_amd import fem
    import inspect
    src = inspect.getsource(fem.FEMSolver.solve_linear_system)
    assert "Unknown solver method" in src
