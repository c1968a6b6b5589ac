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


def test_which_sigma_selection_matches_eigsh():
    """_select picks what scipy's eigsh returns for which / sigma, ascending, on an
    indefinite pencil (checked against ARPACK on the CPU, the reference's call
    eigen_solver.pyx:225-235)."""
    import scipy.linalg as sla
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from qdsim_amd.eigen import _select
    n = 60
    rng = np.random.default_rng(3)
    A = sp.diags([-np.ones(n - 1), np.linspace(-3, 2, n), -np.ones(n - 1)], [-1, 0, 1]).tocsc()
    B = sp.diags([0.1 * np.ones(n - 1), 1 + rng.random(n), 0.1 * np.ones(n - 1)], [-1, 0, 1]).tocsc()
    w = sla.eigh(A.toarray(), B.toarray(), eigvals_only=True)
    assert w.min() < 0 < w.max()
    k = 5
    for which, sigma in (("SA", None), ("LA", None), ("SM", None), ("LM", None), ("LM", 0.3), ("LM", -1.7)):
        ref = spla.eigsh(A, k, M=B, which=which, sigma=sigma, tol=1e-12, return_eigenvectors=False)
        got = w[_select(w, k, which, sigma)]
        assert np.allclose(got, np.sort(ref), rtol=0, atol=1e-9), (which, sigma, got, np.sort(ref))
        assert np.all(np.diff(got) >= 0)
    with pytest.raises(ValueError):
        _select(w, k, "SM", 0.3)
    with pytest.raises(ValueError):
        _select(w, k, "XX", None)


def test_default_block_never_below_k():
    from qdsim_amd.schrodinger import default_block
    for n in range(4, 400, 7):
        for k in range(1, 40):
            b = default_block(k, n)
            if b is None:
                # no multiple-of-4 block b >= k with 3 b <= n exists
                assert 3 * max(4, (k + 3) // 4 * 4) > n
            else:
                assert b >= k and b % 4 == 0 and 3 * b <= n
    assert default_block(10, 40) == 12      # was 8 < k
    assert default_block(16, 2_356_225) == 24


def test_breakdown_reason():
    from qdsim_amd.schrodinger import breakdown_reason
    assert breakdown_reason(np.array([-0.1, 0.2]), True) is None
    assert breakdown_reason(np.array([-0.1, 0.2]), False) == "not converged"
    assert breakdown_reason(np.array([np.nan, 0.2]), True) == "non-finite Ritz values"
    assert breakdown_reason(None, False, np.linalg.LinAlgError("x")).startswith("Rayleigh-Ritz failed")


def test_degenerate_clusters():
    from qdsim_amd.adjoint import degenerate_clusters
    lam = np.array([-0.172, -0.0703, -0.0703 + 1e-12, -0.0083, 0.01, 0.01, 0.01])
    assert degenerate_clusters(lam) == [[1, 2], [4, 5, 6]]
    assert degenerate_clusters(np.array([-0.172, -0.07028, -0.07002])) == []
