"""Adaptive refinement: indicator on the GPU, conforming red-green refinement on
the host, higher-order DOFs regenerated for the refined triangulation.

Reference: FEMSolver::adapt_mesh (backend/src/fem.cpp:486-542) flags the elements
whose least-squares gradient of an eigenvector exceeds a threshold
(AdaptiveMesh::computeRefinementFlags, backend/src/adaptive_mesh.cpp:431-475),
red-refines them (AdaptiveMesh::refineMesh, :61-230) and reassembles.  The
reference splits each flagged triangle into four and keeps its neighbours as
they are, so the refined mesh has hanging nodes, and its P2/P3 children reuse
placeholder edge nodes (:145-186).  Here the flagged triangles are refined red
(four children), a triangle with two or three refined edges becomes red too
(repeated to a fixed point), a triangle with one refined edge is bisected green
(two children), so the result is conforming; P2/P3 DOFs are then generated from
the refined vertex mesh with the element-local order of the element tables
(edge DOFs numbered once per edge, the first one nearer the edge's lower vertex).
The indicator is qds_gradient_indicator; the refinement is host work on index
arrays, followed by the usual host pattern build of the new mesh.
"""
import numpy as np
import torch

from . import _lib
from .device import ptr, stream_ptr

_NPE = {1: 3, 2: 6, 3: 10}


def gradient_indicator(conn, x_nm, y_nm, psi, npe):
    """|least-squares gradient| of the nodal field psi per element, in 1/m (device)."""
    dev = conn.device
    psi = torch.as_tensor(np.asarray(psi, dtype=np.float64) if not torch.is_tensor(psi) else psi,
                          dtype=torch.float64, device=dev).contiguous()
    out = torch.empty(conn.shape[0], dtype=torch.float64, device=dev)
    _lib.check(_lib.lib.qds_gradient_indicator(int(npe), int(conn.shape[0]), ptr(conn), ptr(x_nm), ptr(y_nm),
                                               ptr(psi), 1e-9, ptr(out), stream_ptr()), "qds_gradient_indicator")
    return out


def _edges(tri):
    """Unique undirected edges of a triangulation and each triangle's edge ids in
    local order (01, 12, 20)."""
    loc = np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], axis=1)      # (E, 3, 2)
    key = np.sort(loc, axis=2).reshape(-1, 2)
    edges, inv = np.unique(key, axis=0, return_inverse=True)
    return edges, inv.reshape(-1, 3)


def refine_red_green(x, y, tri, flags):
    """Conforming refinement of a P1 triangulation (vertex coordinates x, y,
    connectivity tri (E, 3), boolean flags (E,)).  Returns (x, y, tri, parent)
    with parent[c] the old triangle of each new one."""
    tri = np.asarray(tri, dtype=np.int64)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    edges, te = _edges(tri)
    marked = np.zeros(len(edges), dtype=bool)
    marked[te[np.asarray(flags, dtype=bool)].ravel()] = True
    while True:                        # closure: two or more refined edges -> red
        cnt = marked[te].sum(axis=1)
        grow = (cnt >= 2) & ~marked[te].all(axis=1)
        if not grow.any():
            break
        marked[te[grow].ravel()] = True
    mid = np.full(len(edges), -1, dtype=np.int64)
    nm = int(marked.sum())
    mid[marked] = len(x) + np.arange(nm)
    e = edges[marked]
    x2 = np.concatenate([x, 0.5 * (x[e[:, 0]] + x[e[:, 1]])])
    y2 = np.concatenate([y, 0.5 * (y[e[:, 0]] + y[e[:, 1]])])
    m = mid[te]                        # (E, 3) midpoint of edge 01, 12, 20 or -1
    cnt = (m >= 0).sum(axis=1)
    out, parent = [], []
    keep = np.nonzero(cnt == 0)[0]
    out.append(tri[keep])
    parent.append(keep)
    red = np.nonzero(cnt == 3)[0]
    if red.size:
        v0, v1, v2 = tri[red, 0], tri[red, 1], tri[red, 2]
        m01, m12, m20 = m[red, 0], m[red, 1], m[red, 2]
        for c in ([v0, m01, m20], [m01, v1, m12], [m20, m12, v2], [m01, m12, m20]):
            out.append(np.stack(c, axis=1))
            parent.append(red)
    green = np.nonzero(cnt == 1)[0]
    if green.size:
        k = np.argmax(m[green] >= 0, axis=1)          # local edge that is split
        a = tri[green, k]                              # edge (a, b), opposite vertex c
        b = tri[green, (k + 1) % 3]
        c = tri[green, (k + 2) % 3]
        mm = m[green, k]
        out.append(np.stack([a, mm, c], axis=1))
        parent.append(green)
        out.append(np.stack([mm, b, c], axis=1))
        parent.append(green)
    return x2, y2, np.concatenate(out).astype(np.int64), np.concatenate(parent)


def elevate(x, y, tri, order):
    """Nodes and connectivity of the P_order space on a P1 triangulation: vertices
    first, then the edge DOFs (one per edge for P2, two for P3, the first nearer
    the edge's lower vertex), then the P3 bubbles; local order as the element
    tables (v0 v1 v2, edges 01 12 20, bubble)."""
    tri = np.asarray(tri, dtype=np.int64)
    if order == 1:
        return np.asarray(x, float), np.asarray(y, float), tri
    edges, te = _edges(tri)
    nv, ne = len(x), len(edges)
    xa, ya = x[edges[:, 0]], y[edges[:, 0]]
    xb, yb = x[edges[:, 1]], y[edges[:, 1]]
    loc = [(0, 1), (1, 2), (2, 0)]
    if order == 2:
        X = np.concatenate([x, 0.5 * (xa + xb)])
        Y = np.concatenate([y, 0.5 * (ya + yb)])
        conn = np.concatenate([tri, nv + te], axis=1)
        return X, Y, conn
    t = 1.0 / 3.0
    X = np.concatenate([x, (1 - t) * xa + t * xb, t * xa + (2 * t) * xb])
    Y = np.concatenate([y, (1 - t) * ya + t * yb, t * ya + (2 * t) * yb])
    near_lo, near_hi = nv + np.arange(ne), nv + ne + np.arange(ne)    # DOFs at 2/3 lo + 1/3 hi, 1/3 lo + 2/3 hi
    cols = [tri[:, 0], tri[:, 1], tri[:, 2]]
    for j, (i0, i1) in enumerate(loc):
        fwd = tri[:, i0] < tri[:, i1]                  # local start vertex is the edge's lower vertex
        eid = te[:, j]
        cols.append(np.where(fwd, near_lo[eid], near_hi[eid]))   # DOF nearer the local start vertex
        cols.append(np.where(fwd, near_hi[eid], near_lo[eid]))
    nb = len(tri)
    X = np.concatenate([X, x[tri].mean(axis=1)])
    Y = np.concatenate([Y, y[tri].mean(axis=1)])
    cols.append(nv + 2 * ne + np.arange(nb))
    return X, Y, np.stack(cols, axis=1)


class TriMesh:
    """Unstructured triangulation (SI coordinates) with the accessors of the FEM
    assembly (like FileMesh); keeps its P1 vertex triangulation for further
    refinement."""

    def __init__(self, vx, vy, vtri, order=1):
        self.vx = np.ascontiguousarray(vx, dtype=np.float64)
        self.vy = np.ascontiguousarray(vy, dtype=np.float64)
        self.vtri = np.ascontiguousarray(vtri, dtype=np.int64)
        self.element_order = int(order)
        X, Y, conn = elevate(self.vx, self.vy, self.vtri, self.element_order)
        self._x, self._y, self._conn = X, Y, conn
        self.num_nodes = int(X.size)
        self.num_elements = int(conn.shape[0])
        self.nodes_per_element = int(conn.shape[1])
        self.Lx = float(X.max() - X.min())
        self.Ly = float(Y.max() - Y.min())
        self.nx = self.ny = None

    def get_nodes(self):
        return self._x.copy(), self._y.copy()

    def get_elements(self):
        return self._conn.astype(np.int32)

    def get_vertex_elements(self):
        return self.vtri.astype(np.int32)


def vertex_mesh(mesh):
    """(x, y, tri) of the P1 vertex triangulation under any mesh of this package."""
    if isinstance(mesh, TriMesh):
        return mesh.vx, mesh.vy, mesh.vtri
    x, y = mesh.get_nodes()
    tri = np.asarray(mesh.get_vertex_elements() if hasattr(mesh, "get_vertex_elements")
                     else mesh.get_elements()[:, :3], dtype=np.int64)
    used = np.unique(tri)
    remap = np.full(len(x), -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    return np.asarray(x)[used], np.asarray(y)[used], remap[tri]


def refine_mesh(mesh, flags):
    """Conforming refinement of ``mesh`` (any order) on the flagged elements;
    returns a TriMesh of the same order."""
    vx, vy, vtri = vertex_mesh(mesh)
    x2, y2, t2, _ = refine_red_green(vx, vy, vtri, flags)
    return TriMesh(x2, y2, t2, getattr(mesh, "element_order", 1))
