"""Adaptive refinement on the GPU: the indicator kernel against the oracle, and
FEMSolver.adapt_mesh end to end (conforming refinement, reassembly checked
against the oracle assembly of the refined mesh, mesh-file cache)."""
import numpy as np
import pytest

from oracle import adapt_ref, fem_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2, 3])
def test_gradient_indicator_matches_oracle(order):
    import torch
    from qdsim_amd.adapt import TriMesh, gradient_indicator, refine_red_green, vertex_mesh
    from qdsim_amd.fem import FEMSpace
    from qdsim_amd.mesh import SimpleMesh
    m = SimpleMesh(9, 7, 60e-9, 45e-9, 1)
    x, y, t = vertex_mesh(m)
    x2, y2, t2, _ = refine_red_green(x, y, t, np.arange(len(t)) % 4 == 0)
    tm = TriMesh(x2, y2, t2, order)
    sp_ = FEMSpace(tm)
    X, Y = tm.get_nodes()
    psi = np.sin(X / 7e-9) * np.cos(Y / 9e-9) + np.random.default_rng(order).standard_normal(len(X)) * 0.01
    print(f"indicator P{order}: {tm.num_elements} elements", flush=True)
    g = gradient_indicator(sp_.conn, sp_.x, sp_.y, psi, sp_.npe).cpu().numpy()
    ref = adapt_ref.gradient_norms(X, Y, tm.get_elements(), psi)
    assert np.allclose(g, ref, rtol=1e-9, atol=1e-9 * ref.max())


@pytest.mark.parametrize("order", [1, 3])
def test_adapt_mesh_other_orders(order):
    """P1 and P3 meshes: refinement conforming, reassembled mass sums to the area."""
    from qdsim_amd.fem import FEMSolver
    from qdsim_amd.mesh import SimpleMesh
    m = SimpleMesh(13, 9, 60e-9, 40e-9, order)
    s = FEMSolver(m)
    s.assemble_mass_matrix()
    x, y = m.get_nodes()
    psi = np.tanh((x - 30e-9) / 3e-9)
    ref = adapt_ref.gradient_norms(x, y, m.get_elements(), psi)
    thr = float(np.quantile(ref, 0.75)) * (1 + 1e-6)
    flags = s.adapt_mesh(psi, thr)
    assert flags.sum() > 0 and s.mesh.num_elements > m.num_elements
    assert np.isclose(s.mass_matrix.sum(), 60e-9 * 40e-9, rtol=1e-10)
    X, Y = s.mesh.get_nodes()
    conn = s.mesh.get_elements().astype(np.int64)
    _, _, M = fem_ref.element_matrices(order, X[conn[:, :3]], Y[conn[:, :3]], 2 * order + 2)
    Mref = fem_ref.assemble(conn, s.mesh.num_nodes, M)
    assert abs(s.mass_matrix - Mref).max() <= 1e-10 * abs(Mref).max()


def test_adapt_mesh_refines_reassembles_and_caches(tmp_path):
    from qdsim_amd.adapt import gradient_indicator
    from qdsim_amd.fem import FEMSolver
    from qdsim_amd.mesh import SimpleMesh
    Lx, Ly = 80e-9, 60e-9
    m = SimpleMesh(17, 13, Lx, Ly, 2)
    s = FEMSolver(m)
    s.assemble_stiffness_matrix()
    s.assemble_mass_matrix()
    x, y = m.get_nodes()
    psi = np.exp(-((x - 40e-9) ** 2 + (y - 30e-9) ** 2) / (2 * (8e-9) ** 2))
    ref = adapt_ref.gradient_norms(x, y, m.get_elements(), psi)
    # symmetric field: many elements share a value; keep the threshold off them so
    # rounding-level differences between kernel and oracle cannot flip a flag
    thr = float(np.quantile(ref, 0.8)) * (1 + 1e-6)
    assert np.min(np.abs(ref - thr)) > 1e-9 * thr
    flags = s.adapt_mesh(psi, thr, cache_dir=str(tmp_path))
    assert np.array_equal(flags, ref > thr)
    new = s.mesh
    print(f"adapt: {m.num_elements} -> {new.num_elements} elements, {m.num_nodes} -> {new.num_nodes} nodes")
    assert new.num_elements >= m.num_elements + 3 * int(flags.sum())
    X, Y = new.get_nodes()
    conn = new.get_elements().astype(np.int64)
    K, _, M = fem_ref.element_matrices(2, X[conn[:, :3]], Y[conn[:, :3]], 6)
    Kref = fem_ref.assemble(conn, new.num_nodes, K)
    Mref = fem_ref.assemble(conn, new.num_nodes, M)
    assert abs(s.stiffness_matrix - Kref).max() <= 1e-10 * abs(Kref).max()
    assert abs(s.mass_matrix - Mref).max() <= 1e-10 * abs(Mref).max()
    assert np.isclose(s.mass_matrix.sum(), Lx * Ly, rtol=1e-10)
    assert abs(s.stiffness_matrix @ np.ones(new.num_nodes)).max() <= 1e-10 * abs(Kref).max()
    # the same eigenvector norm loads the cached mesh file
    s2 = FEMSolver(m)
    s2.adapt_mesh(psi, thr, cache_dir=str(tmp_path))
    assert s2.mesh.num_nodes == new.num_nodes and s2.mesh.num_elements == new.num_elements
    # a second adaptation refines the refined mesh further
    X2, Y2 = s.mesh.get_nodes()
    psi2 = np.exp(-((X2 - 40e-9) ** 2 + (Y2 - 30e-9) ** 2) / (2 * (8e-9) ** 2))
    g2 = gradient_indicator(s.space.conn, s.space.x, s.space.y, psi2, s.space.npe).cpu().numpy()
    n_before = s.mesh.num_elements
    s.adapt_mesh(psi2, float(np.quantile(g2, 0.9)))
    assert s.mesh.num_elements > n_before
