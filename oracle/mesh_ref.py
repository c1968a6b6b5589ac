"""Loop-based restatement of the reference mesh numbering (oracle; test use only).

* P1: qdsim_cython/core/mesh_minimal.pyx:53-89 (vertex id j*nx+i, triangles
  [n0,n1,n2], [n1,n3,n2]).
* P2: backend/src/mesh.cpp:66-115 -- midpoints created through a first-appearance
  cache keyed by std::minmax(edge), call order m01, m12, m20, m23, m13, m03.
* P3: backend/src/mesh.cpp:117-154 with the edge-duplication placeholder replaced
  by two DOFs per edge, assigned consecutively at first appearance, the first one
  nearer the smaller vertex id; centroids c012, c132 after the P2 midpoints.

The cache is a Python dict, exactly like the std::map it restates; this is
O(E) Python work and is meant for small meshes.
"""
import numpy as np


def simple_mesh_p1(nx, ny, Lx, Ly):
    """mesh_minimal.pyx:53-89.  Returns (x, y, elements[E,3])."""
    dx = Lx / (nx - 1)
    dy = Ly / (ny - 1)
    xs, ys = [], []
    for j in range(ny):
        for i in range(nx):
            xs.append(i * dx)
            ys.append(j * dy)
    el = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            n0 = j * nx + i
            n1 = j * nx + (i + 1)
            n2 = (j + 1) * nx + i
            n3 = (j + 1) * nx + (i + 1)
            el.append([n0, n1, n2])
            el.append([n1, n3, n2])
    return np.array(xs), np.array(ys), np.array(el, dtype=np.int64).reshape(-1, 3)


def _vertices(nx, ny, Lx, Ly):
    dx = Lx / (nx - 1)
    dy = Ly / (ny - 1)
    pts = []
    for j in range(ny):
        for i in range(nx):
            pts.append((i * dx, j * dy))
    return pts


def mesh_cpp_p2(nx, ny, Lx, Ly, orphan_nodes=True):
    """mesh.cpp:66-115 with vertex counts (nx, ny).  Returns (x, y, elements[E,6])."""
    pts = _vertices(nx, ny, Lx, Ly)
    cache = {}

    def mid(a, b):
        key = (min(a, b), max(a, b))
        if key not in cache:
            pa, pb = pts[a], pts[b]
            pts.append((0.5 * (pa[0] + pb[0]), 0.5 * (pa[1] + pb[1])))
            cache[key] = len(pts) - 1
        return cache[key]

    el = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            n0 = j * nx + i
            n1 = n0 + 1
            n2 = n0 + nx
            n3 = n2 + 1
            m01 = mid(n0, n1)
            m12 = mid(n1, n2)
            m20 = mid(n2, n0)
            m23 = mid(n2, n3)
            m13 = mid(n1, n3)
            if orphan_nodes:
                mid(n0, n3)
            el.append([n0, n1, n2, m01, m12, m20])
            el.append([n1, n3, n2, m13, m23, m12])
    p = np.array(pts)
    return p[:, 0].copy(), p[:, 1].copy(), np.array(el, dtype=np.int64)


def mesh_p3(nx, ny, Lx, Ly):
    """mesh.cpp:66-154 with two DOFs per edge.  Returns (x, y, elements[E,10])."""
    pts = _vertices(nx, ny, Lx, Ly)
    cache = {}
    cent = {}

    def edge(a, b):
        lo, hi = min(a, b), max(a, b)
        if (lo, hi) not in cache:
            pl, ph = pts[lo], pts[hi]
            pts.append(((2 * pl[0] + ph[0]) / 3, (2 * pl[1] + ph[1]) / 3))
            pts.append(((pl[0] + 2 * ph[0]) / 3, (pl[1] + 2 * ph[1]) / 3))
            cache[(lo, hi)] = len(pts) - 2
        first = cache[(lo, hi)]
        # (dof nearer a, dof nearer b)
        return (first, first + 1) if a == lo else (first + 1, first)

    def centroid(a, b, c):
        key = tuple(sorted((a, b, c)))
        if key not in cent:
            pts.append(((pts[a][0] + pts[b][0] + pts[c][0]) / 3, (pts[a][1] + pts[b][1] + pts[c][1]) / 3))
            cent[key] = len(pts) - 1
        return cent[key]

    el = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            n0 = j * nx + i
            n1 = n0 + 1
            n2 = n0 + nx
            n3 = n2 + 1
            # P2 pass of mesh.cpp:106-111 creates the edges in this order
            edge(n0, n1); edge(n1, n2); edge(n2, n0); edge(n2, n3); edge(n1, n3)
            c012 = centroid(n0, n1, n2)
            c132 = centroid(n1, n3, n2)
            a01, b01 = edge(n0, n1)
            a12, b12 = edge(n1, n2)
            a20, b20 = edge(n2, n0)
            el.append([n0, n1, n2, a01, b01, a12, b12, a20, b20, c012])
            a13, b13 = edge(n1, n3)
            a32, b32 = edge(n3, n2)
            a21, b21 = edge(n2, n1)
            el.append([n1, n3, n2, a13, b13, a32, b32, a21, b21, c132])
    p = np.array(pts)
    return p[:, 0].copy(), p[:, 1].copy(), np.array(el, dtype=np.int64)


def reference_mesh(nx, ny, Lx, Ly, order, orphan_nodes=False):
    if order == 1:
        return simple_mesh_p1(nx, ny, Lx, Ly)
    if order == 2:
        return mesh_cpp_p2(nx, ny, Lx, Ly, orphan_nodes)
    return mesh_p3(nx, ny, Lx, Ly)


def mesh_file_text(x, y, elements, quadratic=None, cubic=None, order=1):
    """backend/src/mesh.cpp:258-289 (Mesh::save): header ``nodes elements order``,
    ``x y`` per node, three vertex ids per element, then the quadratic block (order
    >= 2) and the cubic block (order 3), one element per line.  Doubles go through
    a default std::ostream: printf's %g (precision 6)."""
    out = [f"{len(x)} {len(elements)} {order}"]
    for a, b in zip(x, y):
        out.append("%g %g" % (a, b))
    for e in elements:
        out.append(" ".join(str(int(v)) for v in e[:3]))
    if order >= 2:
        for e in quadratic:
            out.append(" ".join(str(int(v)) for v in e))
    if order == 3:
        for e in cubic:
            out.append(" ".join(str(int(v)) for v in e))
    return "\n".join(out) + "\n"
