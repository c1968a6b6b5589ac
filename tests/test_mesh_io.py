"""Mesh files in the reference text format, CSR interchange, pattern cache (CPU)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import mesh_ref
from qdsim_amd.mesh import SimpleMesh
from qdsim_amd.mesh_io import (FileMesh, load_csr, load_mesh, load_pattern, save_csr, save_mesh,
                               save_pattern)


def test_p1_file_matches_reference_generator_and_writer(tmp_path):
    """A P1 mesh file equals, byte for byte, what Mesh::save writes for the mesh of
    Mesh::generateTriangularMesh (restated in the oracle)."""
    for nx, ny, Lx, Ly in ((4, 3, 25e-9, 20e-9), (7, 5, 1e-7, 3e-8)):
        x, y, el = mesh_ref.simple_mesh_p1(nx, ny, Lx, Ly)
        want = mesh_ref.mesh_file_text(x, y, el, order=1)
        fn = tmp_path / "m.txt"
        save_mesh(SimpleMesh(nx, ny, Lx, Ly, 1), fn)
        assert fn.read_text() == want


def test_p2_file_matches_reference_numbering(tmp_path):
    """P2 with the reference's orphan midpoints: ids and text identical to the
    restated generator (mesh.cpp:66-115) written by the restated Mesh::save."""
    x, y, el = mesh_ref.mesh_cpp_p2(5, 4, 40e-9, 30e-9, orphan_nodes=True)
    want = mesh_ref.mesh_file_text(x, y, el, quadratic=el, order=2)
    fn = tmp_path / "m2.txt"
    save_mesh(SimpleMesh(5, 4, 40e-9, 30e-9, 2, orphan_nodes=True), fn)
    assert fn.read_text() == want


def test_known_file_text(tmp_path):
    """Hand-computed file of a 2 x 1-cell P1 mesh, 10 nm x 5 nm: the default
    ostream prints 5e-09, 1e-08, 0."""
    fn = tmp_path / "k.txt"
    save_mesh(SimpleMesh(3, 2, 10e-9, 5e-9, 1), fn)
    assert fn.read_text() == ("6 4 1\n0 0\n5e-09 0\n1e-08 0\n0 5e-09\n5e-09 5e-09\n1e-08 5e-09\n"
                              "0 1 3\n1 4 3\n1 2 4\n2 5 4\n")
    m = load_mesh(fn)
    assert isinstance(m, FileMesh) and m.element_order == 1 and m.num_nodes == 6
    assert np.array_equal(m.get_elements(), [[0, 1, 3], [1, 4, 3], [1, 2, 4], [2, 5, 4]])
    assert np.array_equal(m.get_nodes()[0], [0, 5e-9, 1e-8, 0, 5e-9, 1e-8])


def test_load_reference_style_p2_file(tmp_path):
    """A P2 file as the reference writes it (vertex block, then the 6-node block)
    loads with the 6-node connectivity."""
    text = ("6 2 2\n0 0\n1 0\n0 1\n0.5 0\n0.5 0.5\n0 0.5\n0 1 2\n1 2 0\n"
            "0 1 2 3 4 5\n1 2 0 4 5 3\n")
    fn = tmp_path / "p2.txt"
    fn.write_text(text)
    m = load_mesh(fn)
    assert m.element_order == 2 and m.get_elements().shape == (2, 6)
    assert np.array_equal(m.get_vertex_elements(), [[0, 1, 2], [1, 2, 0]])
    assert np.array_equal(m.get_elements()[1], [1, 2, 0, 4, 5, 3])


@pytest.mark.parametrize("order", [1, 2, 3])
def test_round_trip(tmp_path, order):
    m = SimpleMesh(6, 5, 50e-9, 37e-9, order)
    fn = tmp_path / f"rt{order}.txt"
    m.save(fn)
    r = SimpleMesh.load(fn)
    assert np.array_equal(r.get_elements(), m.get_elements())
    assert np.array_equal(r.get_vertex_elements(), m.get_vertex_elements())
    for a, b in zip(r.get_nodes(), m.get_nodes()):
        assert np.allclose(a, b, rtol=5e-6, atol=0)   # six significant digits


def test_bad_files(tmp_path):
    fn = tmp_path / "bad.txt"
    fn.write_text("3 1 4\n")
    with pytest.raises(ValueError):
        load_mesh(fn)
    fn.write_text("3 1 1\n0 0\n1 0\n")
    with pytest.raises(ValueError):
        load_mesh(fn)
    fn.write_text("3 1 1\n0 0\n1 0\n0 1\n0 1 3\n")
    with pytest.raises(ValueError):
        load_mesh(fn)


def test_csr_interchange_and_pattern_cache(tmp_path):
    from qdsim_amd import kernels
    rng = np.random.default_rng(3)
    A = sp.random(40, 40, density=0.1, random_state=rng, format="csr")
    save_csr(tmp_path / "a.npz", A)
    B = load_csr(tmp_path / "a.npz")
    assert (B != A).nnz == 0 and B.format == "csr"
    assert (sp.load_npz(tmp_path / "a.npz") != A).nnz == 0     # plain scipy reads it
    m = SimpleMesh(9, 7, 1e-7, 1e-7, 2)
    pat = kernels.pattern_build(m.get_elements(), m.num_nodes)
    save_pattern(tmp_path / "p.npz", pat, m.num_nodes, 6)
    q = load_pattern(tmp_path / "p.npz", n_dof=m.num_nodes, npe=6)
    for k in ("indptr", "indices", "seg_ptr", "seg_src"):
        assert np.array_equal(q[k], pat[k])
    assert q["nnz"] == pat["nnz"]
    with pytest.raises(ValueError):
        load_pattern(tmp_path / "p.npz", n_dof=m.num_nodes + 1)
