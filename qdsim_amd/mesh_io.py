"""Files on either side of the FEM path: the reference's mesh text format, the
scipy CSR interchange format, and a cache of the host-built pattern and scatter
maps.

Mesh text format (backend/src/mesh.cpp:258-289 ``Mesh::save``, :305-339
``Mesh::load``): a header line ``num_nodes num_elements element_order``, one line
``x y`` per node, one line of three vertex ids per element, then for order >= 2
one line of six ids per element (the quadratic connectivity) and for order 3 one
line of ten ids per element (the cubic connectivity).  Numbers are written the
way a default C++ ``std::ostream`` writes them: ``%g`` (six significant digits)
for coordinates, plain integers for ids.  The writer reproduces that byte for
byte, so files move between the two implementations in both directions; like
the reference, a saved mesh keeps only six significant digits of its coordinates.

CSR interchange: ``save_csr`` / ``load_csr`` read and write scipy's ``.npz``
sparse format (``scipy.sparse.save_npz``), the format in which the reference's
Python layer holds its matrices (``scipy.sparse.csr_matrix``,
qdsim_cython/solvers/fem_solver.pyx:142-145).

Pattern cache: ``save_pattern`` / ``load_pattern`` store the CSR pattern of a
mesh's element cliques with the element-to-nonzero scatter map
(``kernels.pattern_build``), so a large mesh can skip the host build;
``FEMSpace(mesh, pattern=...)`` takes the loaded dictionary.
"""
import numpy as np
import scipy.sparse as sp


class FileMesh:
    """Mesh read from a reference mesh file: nodes, connectivity, element order.

    Exposes the accessors the FEM assembly uses (``get_nodes``, ``get_elements``,
    ``num_nodes``, ``element_order``), like ``SimpleMesh``; it has no lattice
    structure, so it serves assembly (``FEMSolver``), not the lattice eigensolver."""

    def __init__(self, x, y, vertex_elements, elements, element_order):
        self._x = np.ascontiguousarray(x, dtype=np.float64)
        self._y = np.ascontiguousarray(y, dtype=np.float64)
        self._tri = np.ascontiguousarray(vertex_elements, dtype=np.int64)
        self._conn = np.ascontiguousarray(elements, dtype=np.int64)
        self.element_order = int(element_order)
        self.num_nodes = int(self._x.size)
        self.num_elements = int(self._conn.shape[0])
        self.nodes_per_element = int(self._conn.shape[1])
        # extent of the node cloud; a file carries no grid dimensions
        self.Lx = float(self._x.max() - self._x.min()) if self.num_nodes else 0.0
        self.Ly = float(self._y.max() - self._y.min()) if self.num_nodes else 0.0
        self.nx = self.ny = None

    def get_nodes(self):
        return self._x.copy(), self._y.copy()

    def get_elements(self):
        return self._conn.astype(np.int32)

    def get_vertex_elements(self):
        return self._tri.astype(np.int32)


_NPE = {1: 3, 2: 6, 3: 10}


def _fmt_g(v):
    # std::ostream's default floating-point output: %g with precision 6
    return "%g" % v


def save_mesh(mesh, filename):
    """Write ``mesh`` in the reference text format (mesh.cpp:258-289)."""
    x, y = mesh.get_nodes()
    conn = np.asarray(mesh.get_elements())
    p = int(getattr(mesh, "element_order", 1))
    if p not in _NPE or conn.shape[1] != _NPE[p]:
        raise ValueError(f"connectivity with {conn.shape[1]} nodes per element does not match order {p}")
    tri = conn[:, :3]
    lines = [f"{len(x)} {conn.shape[0]} {p}"]
    lines += [f"{_fmt_g(a)} {_fmt_g(b)}" for a, b in zip(x.tolist(), y.tolist())]
    lines += [" ".join(map(str, row)) for row in tri.tolist()]
    if p == 2:
        lines += [" ".join(map(str, row)) for row in conn.tolist()]
    if p == 3:
        # the reference's P3 generator also fills a quadratic block (vertices and
        # edge midpoints, which are its cubic edge nodes: mesh.cpp:106-149); this
        # P3 mesh has two nodes per edge and no midpoints, so the quadratic block
        # lists the first node of each edge (01, 12, 20); readers of a cubic file
        # assemble from the cubic block
        quad = conn[:, [0, 1, 2, 3, 5, 7]]
        lines += [" ".join(map(str, row)) for row in quad.tolist()]
        lines += [" ".join(map(str, row)) for row in conn.tolist()]
    with open(filename, "w") as f:
        f.write("\n".join(lines) + "\n")


def load_mesh(filename):
    """Read a reference mesh file (mesh.cpp:305-339) into a ``FileMesh``."""
    with open(filename) as f:
        tok = f.read().split()
    if len(tok) < 3:
        raise ValueError(f"{filename}: not a mesh file (header missing)")
    nn, ne, p = int(tok[0]), int(tok[1]), int(tok[2])
    if p not in _NPE:
        raise ValueError(f"{filename}: element order {p} not in 1..3")
    need = 3 + 2 * nn + 3 * ne + (6 * ne if p >= 2 else 0) + (10 * ne if p == 3 else 0)
    if len(tok) < need:
        raise ValueError(f"{filename}: truncated ({len(tok)} numbers, {need} expected)")
    pos = 3
    xy = np.array(tok[pos:pos + 2 * nn], dtype=np.float64).reshape(nn, 2)
    pos += 2 * nn
    tri = np.array(tok[pos:pos + 3 * ne], dtype=np.int64).reshape(ne, 3)
    pos += 3 * ne
    conn = tri
    if p >= 2:
        conn = np.array(tok[pos:pos + 6 * ne], dtype=np.int64).reshape(ne, 6)
        pos += 6 * ne
    if p == 3:
        conn = np.array(tok[pos:pos + 10 * ne], dtype=np.int64).reshape(ne, 10)
    if conn.size and (conn.min() < 0 or conn.max() >= nn):
        raise ValueError(f"{filename}: element node id out of range")
    return FileMesh(xy[:, 0], xy[:, 1], tri, conn, p)


def save_csr(filename, A):
    """scipy's .npz sparse format (``scipy.sparse.save_npz``) of a CSR matrix."""
    sp.save_npz(filename, sp.csr_matrix(A), compressed=False)


def load_csr(filename):
    return sp.load_npz(filename).tocsr()


def save_pattern(filename, pattern, n_dof, npe):
    """Store a ``kernels.pattern_build`` result (CSR pattern + scatter map)."""
    np.savez(filename, indptr=pattern["indptr"], indices=pattern["indices"], seg_ptr=pattern["seg_ptr"],
             seg_src=pattern["seg_src"], n_dof=np.int64(n_dof), npe=np.int64(npe))


def load_pattern(filename, n_dof=None, npe=None):
    with np.load(filename, allow_pickle=False) as z:
        out = {k: np.ascontiguousarray(z[k]) for k in ("indptr", "indices", "seg_ptr", "seg_src")}
        nd, ne = int(z["n_dof"]), int(z["npe"])
    if (n_dof is not None and nd != n_dof) or (npe is not None and ne != npe):
        raise ValueError(f"{filename}: pattern of {nd} DOFs / {ne} nodes per element, "
                         f"expected {n_dof} / {npe}")
    out["nnz"] = int(out["indices"].size)
    return out
