"""Product mesh numbering and the native (host) pattern builder vs the oracle."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fem_ref, mesh_ref
from qdsim_amd.mesh import SimpleMesh


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("orphans", [False, True])
@pytest.mark.parametrize("nxy", [(2, 2), (3, 2), (5, 4), (8, 6)])
def test_numbering_matches_oracle(order, orphans, nxy):
    nx, ny = nxy
    m = SimpleMesh(nx, ny, 10e-9, 8e-9, order, orphans)
    x, y, el = mesh_ref.reference_mesh(nx, ny, 10e-9, 8e-9, order, orphans and order == 2)
    assert m.num_nodes == len(x)
    assert np.array_equal(m.get_elements(), el)
    mx, my = m.get_nodes()
    assert np.allclose(mx, x, rtol=0, atol=1e-22) and np.allclose(my, y, rtol=0, atol=1e-22)


def test_simple_mesh_surface():
    m = SimpleMesh(5, 4, 10e-9, 8e-9)
    assert (m.nx, m.ny, m.num_nodes, m.num_elements) == (5, 4, 20, 24)
    assert m.get_elements().shape == (24, 3) and m.get_elements().dtype == np.int32
    assert m.get_node_at(0) == (0.0, 0.0)
    assert m.is_inside_domain(5e-9, 4e-9) and not m.is_inside_domain(-1e-9, 4e-9)
    assert m.find_nearest_node(0.0, 0.0) == 0
    info = m.get_mesh_info()
    assert info['dx'] == 10e-9 / 4 and info['num_elements'] == 24


@pytest.mark.parametrize("order,cells,ndof", [(3, (512, 512), 2362369), (2, (384, 288), 443713)])
def test_large_dof_counts_and_lattice(order, cells, ndof):
    m = SimpleMesh(cells[0] + 1, cells[1] + 1, 1e-7, 1e-7, order)
    assert m.num_nodes == ndof
    li = m.lattice_index()
    assert np.unique(li).size == ndof            # DOFs <-> lattice points one-to-one


@pytest.mark.parametrize("order", [1, 2, 3])
def test_native_pattern_matches_scipy(order):
    from qdsim_amd import kernels
    m = SimpleMesh(7, 6, 6e-9, 5e-9, order)
    el = m.get_elements()
    n = m.num_nodes
    pat = kernels.pattern_build(el, n, with_elem_map=True)
    A = fem_ref.assemble(el.astype(np.int64), n, np.ones((len(el), el.shape[1], el.shape[1])))
    assert np.array_equal(pat["indptr"], A.indptr) and np.array_equal(pat["indices"], A.indices)
    # scatter map: entry (e, i, j) lands on (conn[e,i], conn[e,j])
    emap = pat["elem_map"]
    rows = np.repeat(np.arange(n), np.diff(pat["indptr"]))
    npe = el.shape[1]
    assert np.array_equal(rows[emap], np.repeat(el[:, :, None], npe, 2))
    assert np.array_equal(pat["indices"][emap], np.repeat(el[:, None, :], npe, 1))
    # inverse map: segments sum element entries exactly like COO->CSR
    rng = np.random.default_rng(0)
    Ke = rng.standard_normal((len(el), npe, npe)); Ke = Ke + Ke.transpose(0, 2, 1)
    iu = np.triu_indices(npe)
    packed = Ke[:, iu[0], iu[1]].ravel()
    vals = np.add.reduceat(packed[pat["seg_src"]], pat["seg_ptr"][:-1])
    ref = fem_ref.assemble(el.astype(np.int64), n, Ke)
    assert np.allclose(vals, ref.data, rtol=1e-14, atol=1e-14)


def test_native_pattern_p2_384x288_nnz():
    from qdsim_amd import kernels
    m = SimpleMesh(385, 289, 80e-9, 60e-9, 2)
    pat = kernels.pattern_build(m.get_elements(), m.num_nodes)
    assert pat["nnz"] == 5092609                  # P2 384 x 288 cells: structured-mesh combinatorics


def test_restrict_pattern():
    from qdsim_amd import kernels
    from qdsim_amd.lattice import Lattice
    m = SimpleMesh(6, 5, 5e-9, 4e-9, 3)
    el = m.get_elements()
    pat = kernels.pattern_build(el, m.num_nodes)
    lat = Lattice(m)
    r_ip, r_ix, r_src = kernels.pattern_restrict(pat["indptr"], pat["indices"], lat.new_of_old, lat.n_free)
    A = sp.csr_matrix((np.arange(pat["nnz"], dtype=float) + 1, pat["indices"], pat["indptr"]))
    sub = A[lat.old_of_new][:, lat.old_of_new].tocsr()
    sub.sort_indices()
    assert np.array_equal(r_ip, sub.indptr) and np.array_equal(r_ix, sub.indices)
    assert np.array_equal(r_src + 1, sub.data.astype(np.int64))


def test_capi_exports_every_declared_symbol():
    import ctypes
    from qdsim_amd import _lib
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "qdsim_hip.h")).read()
    declared = set(re.findall(r"\b(qds_[a-z0-9_]+)\s*\(", hdr))
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.exported_symbols())


def test_product_quadrature_tables():
    from qdsim_amd.quadrature import element_tables, triangle_rule
    from math import factorial
    for deg in (4, 6, 8):
        P, W = triangle_rule(deg)
        Po, Wo = fem_ref.rule(deg)
        assert np.allclose(np.sort(W), np.sort(Wo), atol=1e-15)
    for order in (1, 2, 3):
        t = element_tables(order)
        # product tables vs oracle closed-form shape functions at the same points
        l = t["bary"]
        phi = np.stack(fem_ref.shape_values(order, l[:, 0], l[:, 1], l[:, 2]), -1)
        assert np.allclose(t["phi"], phi, atol=1e-13)


def _lattice_pattern(p, ncx, ncy):
    from qdsim_amd import kernels
    from qdsim_amd.lattice import Lattice
    m = SimpleMesh(ncx + 1, ncy + 1, 1e-7, 1e-7, p)
    conn = np.ascontiguousarray(m.get_elements(), dtype=np.int32)
    pat = kernels.pattern_build(conn, int(m.num_nodes))
    lat = Lattice(m)
    r_ip, r_ix, _ = kernels.pattern_restrict(pat["indptr"], pat["indices"], lat.new_of_old, lat.n_free)
    return lat, r_ip, r_ix


@pytest.mark.parametrize("p,ncx,ncy", [(1, 9, 7), (2, 8, 5), (3, 13, 10), (3, 1, 2)])
def test_lattice_stencil_reproduces_csr(p, ncx, ncy):
    """Host stencil builder: every CSR entry lands in exactly one value slot, the
    class stencils are the FE couplings of the structured triangulation, and the
    product evaluated from the stencil form (offsets, zero slots) equals the CSR
    product exactly (same summation order)."""
    from qdsim_amd import kernels
    lat, r_ip, r_ix = _lattice_pattern(p, ncx, ncy)
    st, slot = kernels.lattice_stencil(p, lat.nx_i, lat.ny_i, r_ip, r_ix)
    assert st.order == p and st.n_class == p * p
    assert np.array_equal(np.sort(slot[slot >= 0]), np.arange(len(r_ix)))
    # stencil sizes of the interior classes: P1 7-point, P2 19/9, P3 37/16/10
    want = {1: [7], 2: [19, 9, 9, 9], 3: [37, 16, 16, 16, 10, 16, 16, 16, 10]}[p]
    if ncx >= 4 and ncy >= 4:
        lens = []
        for c in range(p * p):
            offs = {(st.di[c][k], st.dj[c][k]) for k in range(st.len[c])}
            lens.append(len(offs))
        assert lens == want
    rng = np.random.default_rng(p)
    vals = rng.standard_normal(len(r_ix))
    A = sp.csr_matrix((vals, r_ix, r_ip), shape=(lat.n_free, lat.n_free))
    sv = np.where(slot >= 0, vals[np.maximum(slot, 0)], 0.0)
    X = rng.standard_normal((lat.n_free, 3))
    Y = np.zeros_like(X)
    nx, ny = lat.nx_i, lat.ny_i
    for r in range(lat.n_free):
        j, i = divmod(r, nx)
        I, J = i + 1, j + 1
        c = (J % p) * p + I % p
        L = st.len[c]
        cx, cy = I // p, J // p
        base = (cy * st.tiles_x + cx // st.tile) * st.tile_slots + st.base[c] + (cx % st.tile) * 4
        for k in range(L):
            slot = base + (k // 4) * st.tile * 4 + k % 4
            I2, J2 = I + st.di[c][k], J + st.dj[c][k]
            if 1 <= I2 <= nx and 1 <= J2 <= ny:
                Y[r] += sv[slot] * X[(J2 - 1) * nx + I2 - 1]
            else:
                assert sv[slot] == 0.0
    assert np.array_equal(Y, A @ X)


def test_lattice_stencil_rejects_non_lattice_pattern():
    from qdsim_amd import kernels
    from qdsim_amd._lib import QDSError
    ip = np.array([0, 2, 3, 4, 5], dtype=np.int32)
    ix = np.array([0, 3, 1, 2, 3], dtype=np.int32)   # row 0 couples to a point 3 steps away
    with pytest.raises(QDSError):
        kernels.lattice_stencil(1, 4, 1, ip, ix)
