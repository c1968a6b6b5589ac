"""Structured triangular meshes with P1, P2 and P3 degree-of-freedom numbering.

Call surface: ``SimpleMesh(nx, ny, Lx, Ly)`` of qdsim_cython/core/mesh_minimal.pyx:32
(``nx``/``ny`` are vertex counts, vertex id ``j*nx + i``, cell (i, j) split into
``[n0, n1, n2]`` and ``[n1, n3, n2]``; mesh_minimal.pyx:53-89).  Higher orders add
the DOF numbering of backend/src/mesh.cpp:66-157, computed here in closed form
instead of through a ``std::map`` first-appearance cache:

* P2 (exact reference numbering): after the vertices, every cell creates its new
  edge midpoints in the order m01, m12, m20, m23, m13 (and m03, the unused
  n0-n3 midpoint, when ``orphan_nodes=True``; mesh.cpp:106-112).  Element DOFs
  ``[v0, v1, v2, e01, e12, e20]``.
* P3 (reference-derived): the reference gives both DOFs of an edge one node
  (mesh.cpp:130-149).  Here every edge owns two consecutive ids, assigned in the
  same creation order, the first one nearer the smaller vertex id; the cell then
  creates the two interior nodes c012, c132 (mesh.cpp:117-127).  Element DOFs
  ``[v0, v1, v2, e01a, e01b, e12a, e12b, e20a, e20b, c]`` where the "a" DOF of
  edge (vi, vj) sits at lambda_i = 2/3 (fe_interpolator.cpp:506-513).

Every DOF of an order-p mesh lies on the (p*(nx-1)+1) x (p*(ny-1)+1) lattice of the
p-times refined grid; ``lattice_index`` gives that position.  The GPU eigensolver
stores vectors in lattice order, which turns the FEM operator into a banded
matrix and lets the multigrid preconditioner work on plain 2-D arrays.
"""

import numpy as np


def _cell_grid(nx, ny):
    cx, cy = nx - 1, ny - 1
    jj, ii = np.meshgrid(np.arange(cy, dtype=np.int64), np.arange(cx, dtype=np.int64), indexing="ij")
    return cx, cy, ii.ravel(), jj.ravel()


def _number_dofs(nx, ny, order, orphan_nodes):
    """Return (conn[E, npe] int64, lattice (lx, ly) per DOF, num_dofs)."""
    cx, cy, ci, cj = _cell_grid(nx, ny)
    ncell = cx * cy
    nv = nx * ny
    n0 = cj * nx + ci
    n1 = n0 + 1
    n2 = n0 + nx
    n3 = n2 + 1

    # vertex lattice coordinates
    vj, vi = np.divmod(np.arange(nv, dtype=np.int64), nx)
    if order == 1:
        conn = np.empty((2 * ncell, 3), dtype=np.int64)
        conn[0::2] = np.stack([n0, n1, n2], axis=1)
        conn[1::2] = np.stack([n1, n3, n2], axis=1)
        return conn, (vi, vj), nv

    p = order
    per_edge = 1 if order == 2 else 2
    first_row = (cj == 0).astype(np.int64)
    first_col = (ci == 0).astype(np.int64)
    orphan = 1 if (orphan_nodes and order == 2) else 0
    interior = 0 if order == 2 else 2
    # ids created by each cell, in creation order:
    #   [e01 if j==0], e12, [e20 if i==0], e23, e13, [e03 orphan], (c012, c132 for P3)
    count = per_edge * (first_row + 1 + first_col + 1 + 1) + orphan + interior
    base = nv + np.concatenate([[0], np.cumsum(count)[:-1]])
    off = base.copy()
    e01_new = off.copy(); off += per_edge * first_row
    e12 = off.copy(); off += per_edge
    e20_new = off.copy(); off += per_edge * first_col
    e23 = off.copy(); off += per_edge
    e13 = off.copy(); off += per_edge
    e03 = off.copy(); off += orphan
    c012 = off.copy(); c132 = off + 1
    ndof = int(base[-1] + count[-1]) if ncell else nv

    # bottom edge of cell (i, j): created by cell (i, 0) as e01 or by (i, j-1) as e23
    below = np.maximum(cj - 1, 0) * cx + ci
    e_bot = np.where(cj == 0, e01_new, e23[below])
    # left edge: created by cell (0, j) as e20 or by (i-1, j) as e13
    left = cj * cx + np.maximum(ci - 1, 0)
    e_left = np.where(ci == 0, e20_new, e13[left])

    lat_x = np.empty(ndof, dtype=np.int64)
    lat_y = np.empty(ndof, dtype=np.int64)
    lat_x[:nv] = p * vi
    lat_y[:nv] = p * vj
    bx, by = p * ci, p * cj

    def put(ids, dx, dy, mask=None):
        if mask is None:
            lat_x[ids] = bx + dx
            lat_y[ids] = by + dy
        else:
            lat_x[ids[mask]] = (bx + dx)[mask]
            lat_y[ids[mask]] = (by + dy)[mask]

    if order == 2:
        put(e_bot, 1, 0, cj == 0)
        put(e12, 1, 1)
        put(e_left, 0, 1, ci == 0)
        put(e23, 1, 2)
        put(e13, 2, 1)
        if orphan:
            put(e03, 1, 1)       # the n0-n3 midpoint coincides with the n1-n2 one
        conn = np.empty((2 * ncell, 6), dtype=np.int64)
        conn[0::2] = np.stack([n0, n1, n2, e_bot, e12, e_left], axis=1)
        conn[1::2] = np.stack([n1, n3, n2, e13, e23, e12], axis=1)
        return conn, (lat_x, lat_y), ndof

    # P3: edge id +0 sits nearer the smaller vertex id (left / bottom / n1 end)
    m = cj == 0
    put(e_bot, 1, 0, m); put(e_bot + 1, 2, 0, m)
    put(e12, 2, 1); put(e12 + 1, 1, 2)
    m = ci == 0
    put(e_left, 0, 1, m); put(e_left + 1, 0, 2, m)
    put(e23, 1, 3); put(e23 + 1, 2, 3)
    put(e13, 3, 1); put(e13 + 1, 3, 2)
    put(c012, 1, 1)
    put(c132, 2, 2)
    conn = np.empty((2 * ncell, 10), dtype=np.int64)
    # [n0, n1, n2]: e01 = bottom (a near n0 = left), e12 = diagonal (a near n1),
    #               e20 = left edge (a near n2 = top)
    conn[0::2] = np.stack([n0, n1, n2, e_bot, e_bot + 1, e12, e12 + 1,
                           e_left + 1, e_left, c012], axis=1)
    # [n1, n3, n2]: e01 = right edge (a near n1 = bottom), e12 = top edge
    #               (a near n3 = right), e20 = diagonal (a near n2)
    conn[1::2] = np.stack([n1, n3, n2, e13, e13 + 1, e23 + 1, e23,
                           e12 + 1, e12, c132], axis=1)
    return conn, (lat_x, lat_y), ndof


class SimpleMesh:
    """Structured rectangle mesh; see the module docstring for the numbering.

    Parameters follow mesh_minimal.pyx:32 (``nx``, ``ny`` vertex counts, SI lengths),
    plus ``element_order`` (1, 2, 3) and ``orphan_nodes`` (P2 only: keep the unused
    n0-n3 midpoint of every cell so ids match backend/src/mesh.cpp exactly).
    """

    def __init__(self, nx=10, ny=10, Lx=100e-9, Ly=100e-9, element_order=1, orphan_nodes=False):
        nx, ny = int(nx), int(ny)
        if nx < 2 or ny < 2:
            raise ValueError("SimpleMesh needs at least 2 vertices per direction")
        if element_order not in (1, 2, 3):
            raise ValueError("Element order must be 1 (linear), 2 (quadratic), or 3 (cubic)")
        self.nx, self.ny = nx, ny
        self.Lx, self.Ly = float(Lx), float(Ly)
        self.element_order = int(element_order)
        self.orphan_nodes = bool(orphan_nodes) and element_order == 2
        conn, (lx, ly), ndof = _number_dofs(nx, ny, self.element_order, self.orphan_nodes)
        self._conn = conn
        self.lattice_x = lx
        self.lattice_y = ly
        self.num_nodes = int(ndof)
        self.num_vertices = nx * ny
        self.num_elements = 2 * (nx - 1) * (ny - 1)
        self.nodes_per_element = conn.shape[1]
        p = self.element_order
        self.lattice_shape = (p * (ny - 1) + 1, p * (nx - 1) + 1)   # (rows, cols)
        self._x, self._y = self._coordinates()

    # -- geometry -------------------------------------------------------------
    def _coordinates(self):
        dx = self.Lx / (self.nx - 1)
        dy = self.Ly / (self.ny - 1)
        nv = self.num_vertices
        vj, vi = np.divmod(np.arange(nv), self.nx)
        x = np.empty(self.num_nodes)
        y = np.empty(self.num_nodes)
        x[:nv] = vi * dx
        y[:nv] = vj * dy
        if self.num_nodes > nv:
            p = self.element_order
            hx = dx / p
            hy = dy / p
            x[nv:] = self.lattice_x[nv:] * hx
            y[nv:] = self.lattice_y[nv:] * hy
        return x, y

    def get_nodes(self):
        """(x, y) float64 arrays of all DOF coordinates (mesh_minimal.pyx:91-108)."""
        return self._x.copy(), self._y.copy()

    def get_elements(self):
        """int32 connectivity (E, npe); (E, 3) for P1 (mesh_minimal.pyx:110-126)."""
        return self._conn.astype(np.int32)

    def get_vertex_elements(self):
        """int32 (E, 3) vertex triangles, whatever the order."""
        return self._conn[:, :3].astype(np.int32)

    def get_node_at(self, node_id):
        if node_id < 0 or node_id >= self.num_nodes:
            raise IndexError(f"Node ID {node_id} out of range [0, {self.num_nodes-1}]")
        return (float(self._x[node_id]), float(self._y[node_id]))

    def find_nearest_node(self, x, y):
        d = (self._x - x) ** 2 + (self._y - y) ** 2
        return int(np.argmin(d))

    def is_inside_domain(self, x, y):
        return (0 <= x <= self.Lx) and (0 <= y <= self.Ly)

    def boundary_nodes(self):
        """DOF ids on the outer boundary of the rectangle (lattice test, exact)."""
        ry, rx = self.lattice_shape
        lx, ly = self.lattice_x, self.lattice_y
        on = (lx == 0) | (lx == rx - 1) | (ly == 0) | (ly == ry - 1)
        if self.orphan_nodes:
            pass  # orphan midpoints are interior points
        return np.nonzero(on)[0]

    def lattice_index(self):
        """Row-major lattice position of every DOF (rows = y)."""
        return self.lattice_y * self.lattice_shape[1] + self.lattice_x

    def save(self, filename):
        """Write the mesh in the reference text format (backend/src/mesh.cpp:258-289)."""
        from .mesh_io import save_mesh
        save_mesh(self, filename)

    @staticmethod
    def load(filename):
        """Read a reference mesh file (backend/src/mesh.cpp:305-339) as a FileMesh."""
        from .mesh_io import load_mesh
        return load_mesh(filename)

    def get_mesh_info(self):
        return {
            'nx': self.nx,
            'ny': self.ny,
            'Lx': self.Lx,
            'Ly': self.Ly,
            'num_nodes': self.num_nodes,
            'num_elements': self.num_elements,
            'dx': self.Lx / (self.nx - 1),
            'dy': self.Ly / (self.ny - 1),
            'element_order': self.element_order,
        }


def create_simple_mesh(nx=20, ny=10, Lx=100e-9, Ly=50e-9, element_order=1):
    """Factory of mesh_minimal.pyx:205-224."""
    return SimpleMesh(nx, ny, Lx, Ly, element_order)
