"""Adaptive refinement (CPU): the indicator oracle pinned, conforming red-green
refinement, and P2/P3 DOF generation on the refined triangulation."""
import numpy as np

from oracle import adapt_ref
from qdsim_amd.adapt import TriMesh, elevate, refine_red_green, vertex_mesh
from qdsim_amd.mesh import SimpleMesh
from qdsim_amd.quadrature import dof_positions


def _area(x, y, t):
    return 0.5 * ((x[t[:, 1]] - x[t[:, 0]]) * (y[t[:, 2]] - y[t[:, 0]])
                  - (x[t[:, 2]] - x[t[:, 0]]) * (y[t[:, 1]] - y[t[:, 0]]))


def _edge_counts(t):
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, c = np.unique(e, axis=0, return_counts=True)
    return c


def test_indicator_oracle_pinned():
    m = SimpleMesh(6, 5, 50e-9, 40e-9, 2)
    x, y = m.get_nodes()
    conn = m.get_elements()
    g = adapt_ref.gradient_norms(x, y, conn, 3e7 * x - 4e7 * y + 0.2)      # linear: |grad| = 5e7 exactly
    assert np.allclose(g, 5e7, rtol=1e-9)
    rng = np.random.default_rng(0)
    psi = rng.standard_normal(len(x))
    ref = [np.hypot(*np.linalg.lstsq(np.column_stack([np.ones(6), x[c], y[c]]), psi[c], rcond=None)[0][1:])
           for c in conn]
    assert np.allclose(adapt_ref.gradient_norms(x, y, conn, psi), ref, rtol=1e-10)


def test_red_green_refinement_is_conforming():
    m = SimpleMesh(9, 7, 80e-9, 60e-9, 1)
    x, y, t = vertex_mesh(m)
    rng = np.random.default_rng(1)
    flags = rng.random(len(t)) < 0.2
    x2, y2, t2, parent = refine_red_green(x, y, t, flags)
    a0, a2 = _area(x, y, t), _area(x2, y2, t2)
    assert np.all(a0 > 0) and np.all(a2 > 0)                                   # orientation kept
    assert np.isclose(a2.sum(), a0.sum(), rtol=1e-13)                          # area preserved
    assert np.allclose(np.bincount(parent, a2, len(t)), a0, rtol=1e-12)        # children tile parents
    c0, c2 = _edge_counts(t), _edge_counts(t2)
    assert c2.max() == 2                                                       # no edge in 3+ triangles
    # conforming: the boundary (edges in one triangle) only gains midpoints, so its
    # length is unchanged and no interior edge appears once (no hanging node)
    def blen(xx, yy, tt):
        e = np.sort(np.concatenate([tt[:, [0, 1]], tt[:, [1, 2]], tt[:, [2, 0]]]), axis=1)
        u, c = np.unique(e, axis=0, return_counts=True)
        b = u[c == 1]
        return np.hypot(xx[b[:, 0]] - xx[b[:, 1]], yy[b[:, 0]] - yy[b[:, 1]]).sum()
    assert np.isclose(blen(x2, y2, t2), blen(x, y, t), rtol=1e-13)
    # every flagged triangle is split into four (red)
    nchild = np.bincount(parent, minlength=len(t))
    assert np.all(nchild[flags] == 4) and set(np.unique(nchild)) <= {1, 2, 4}


def test_closure_turns_two_refined_edges_red():
    x = np.array([0.0, 1.0, 0.0, 1.0, -1.0])
    y = np.array([0.0, 0.0, 1.0, 1.0, 1.0])
    t = np.array([[0, 1, 2], [1, 3, 2], [0, 2, 4]])
    x2, y2, t2, parent = refine_red_green(x, y, t, np.array([False, True, False]))
    # triangle 1 red splits edge (1,2) shared with triangle 0 -> green; triangle 2 untouched
    assert np.bincount(parent).tolist() == [2, 4, 1]
    x3, y3, t3, p3 = refine_red_green(x, y, t, np.array([False, True, True]))
    # triangle 0 now has two split edges (1,2) and (0,2) -> red
    assert np.bincount(p3).tolist() == [4, 4, 4]


def test_elevated_dofs_sit_at_reference_positions():
    m = SimpleMesh(6, 5, 50e-9, 40e-9, 1)
    x, y, t = vertex_mesh(m)
    x2, y2, t2, _ = refine_red_green(x, y, t, np.arange(len(t)) % 3 == 0)
    for order in (2, 3):
        X, Y, conn = elevate(x2, y2, t2, order)
        pos = dof_positions(order)
        v = conn[:, :3]
        ex = x2[t2[:, 0]][:, None] + pos[None, :, 0] * (x2[t2[:, 1]] - x2[t2[:, 0]])[:, None] \
            + pos[None, :, 1] * (x2[t2[:, 2]] - x2[t2[:, 0]])[:, None]
        ey = y2[t2[:, 0]][:, None] + pos[None, :, 0] * (y2[t2[:, 1]] - y2[t2[:, 0]])[:, None] \
            + pos[None, :, 1] * (y2[t2[:, 2]] - y2[t2[:, 0]])[:, None]
        assert np.array_equal(v, t2)
        assert np.allclose(X[conn], ex, atol=1e-22) and np.allclose(Y[conn], ey, atol=1e-22)
        # every DOF used, shared DOFs are shared (count = vertices + edge DOFs + bubbles)
        ne = len(np.unique(np.sort(np.concatenate([t2[:, [0, 1]], t2[:, [1, 2]], t2[:, [2, 0]]]), axis=1), axis=0))
        assert len(X) == len(x2) + (order - 1) * ne + (len(t2) if order == 3 else 0)
        assert np.array_equal(np.unique(conn), np.arange(len(X)))
    tm = TriMesh(x2, y2, t2, 3)
    assert tm.get_elements().shape == (len(t2), 10) and tm.num_nodes == len(X)
