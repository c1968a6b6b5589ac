"""Device-resident finite-element space and the FEMSolver call surface.

``FEMSpace`` owns everything that depends only on the mesh: connectivity and DOF
coordinates in HBM (nanometres), the CSR pattern with its element->nonzero
scatter map and the inverse segment map (built once by the native pattern
builder), and the quadrature/shape tables.  ``assemble`` runs the two HIP
kernels (element matrices, segmented gather) and returns CSR value arrays in
HBM.

``FEMSolver`` reproduces qdsim_cython/solvers/fem_solver.pyx:24-411
(CythonFEMSolver): same constructor, methods, return types and error behaviour,
with assembly on the GPU.  Coefficients are sampled at the element centroid, as
in fem_solver.pyx:126-128, unless ``coefficient_sampling='quadrature'``.
"""
import time

import numpy as np
import scipy.sparse as sp
import torch

from . import kernels
from .coefficients import evaluate
from .device import device
from .quadrature import element_tables

NM = 1e-9


class FEMSpace:
    """Mesh, pattern and quadrature of one structured FE mesh on the GPU."""

    def __init__(self, mesh, degree=None, pattern=None):
        self.mesh = mesh
        self.order = int(getattr(mesh, "element_order", 1))
        conn = np.ascontiguousarray(mesh.get_elements(), dtype=np.int32)
        self.conn_host = conn
        self.n_elem, self.npe = conn.shape
        self.nsym = self.npe * (self.npe + 1) // 2
        self.n = int(mesh.num_nodes)
        x, y = mesh.get_nodes()
        self.x_host, self.y_host = x, y
        t0 = time.perf_counter()
        # pattern: built on the host, or a cached one (mesh_io.load_pattern)
        if pattern is None:
            pat = kernels.pattern_build(conn, self.n)
        else:
            pat = pattern
            if (pat["indptr"].size != self.n + 1 or pat["seg_src"].size != self.n_elem * self.npe * self.npe
                    or int(pat["indptr"][-1]) != pat["indices"].size):
                raise ValueError("cached pattern does not belong to this mesh")
        self.pattern_time = time.perf_counter() - t0
        self.nnz = pat["nnz"]
        self.indptr_host = pat["indptr"]
        self.indices_host = pat["indices"]
        dev = device()
        self.dev = dev
        self.conn = torch.from_numpy(conn).to(dev)
        self.x = torch.from_numpy(x / NM).to(dev)
        self.y = torch.from_numpy(y / NM).to(dev)
        self.indptr = torch.from_numpy(pat["indptr"]).to(dev)
        self.indices = torch.from_numpy(pat["indices"]).to(dev)
        self.seg_ptr = torch.from_numpy(pat["seg_ptr"]).to(dev)
        self.seg_src = torch.from_numpy(pat["seg_src"]).to(dev)
        tab = element_tables(self.order, degree)
        self.degree = tab["degree"]
        self.nq = len(tab["w"])
        packed = np.concatenate([tab["w"], tab["phi"].ravel(), tab["dxi"].ravel(), tab["deta"].ravel()])
        self.tables = torch.from_numpy(packed).to(dev)
        self.bary = torch.from_numpy(np.ascontiguousarray(tab["bary"])).to(dev)
        self.centroid_bary = torch.full((1, 3), 1.0 / 3.0, dtype=torch.float64, device=dev)

    # -- coefficient sampling points (nm) ----------------------------------------
    def quad_points(self):
        return kernels.quad_points(self.npe, self.conn, self.x, self.y, self.bary)

    def centroids(self):
        xq, yq = kernels.quad_points(self.npe, self.conn, self.x, self.y, self.centroid_bary)
        return xq[:, 0].contiguous(), yq[:, 0].contiguous()

    def sample(self, func, where="quadrature", scale=1.0):
        """Evaluate a user callback f(x_m, y_m) at quadrature points or centroids.

        Returns a float64 device tensor of shape (E, Q) or (E,), multiplied by scale."""
        if where == "quadrature":
            xq, yq = self.quad_points()
        else:
            xq, yq = self.centroids()
        vals = evaluate(func, xq * NM, yq * NM)
        return (vals * scale).contiguous()

    # -- assembly ---------------------------------------------------------------
    def assemble(self, c=None, v=None, r=None, w=None, stiffness_mass=True, mass=False, cap=False):
        """CSR values (FP64, length nnz) of
             out0 = K(c) + M(v)   (if stiffness_mass)
             out1 = M(r)          (r None -> density 1)
             out2 = M(w)          (CAP absorption)
        in the internal length unit (nm).  K is scale free in 2-D; M carries nm^2."""
        e0, e1, e2 = kernels.assemble_elements(self.npe, self.nq, self.conn, self.x, self.y, self.tables,
                                               c=c, v=v, r=r, w=w, out0=stiffness_mass, out1=mass, out2=cap)
        return kernels.assemble_gather(self.nnz, self.seg_ptr, self.seg_src, e0, e1, e2)

    def element_matrices(self, c=None, v=None, r=None, w=None, which=0):
        e = kernels.assemble_elements(self.npe, self.nq, self.conn, self.x, self.y, self.tables,
                                      c=c, v=v, r=r, w=w, out0=which == 0, out1=which == 1, out2=which == 2)
        return e[which]

    def to_scipy(self, values, scale=1.0):
        data = values.detach().to("cpu").numpy() * scale
        return sp.csr_matrix((data, self.indices_host.copy(), self.indptr_host.copy()),
                             shape=(self.n, self.n))

    def boundary_dofs(self):
        return self.mesh.boundary_nodes()


def _centroid_or_quadrature(space, func, sampling, scale=1.0):
    if func is None:
        return None
    return space.sample(func, "centroid" if sampling == "centroid" else "quadrature", scale)


class FEMSolver:
    """Drop-in for CythonFEMSolver (qdsim_cython/solvers/fem_solver.pyx:24)."""

    def __init__(self, mesh, element_order=1, coefficient_sampling="centroid"):
        self.mesh = mesh
        self._last_funcs = {}
        self.num_nodes = mesh.num_nodes
        self.num_elements = mesh.num_elements
        self.Lx = mesh.Lx
        self.Ly = mesh.Ly
        self.nx = mesh.nx
        self.ny = mesh.ny
        self.element_order = int(getattr(mesh, "element_order", element_order))
        self.coefficient_sampling = coefficient_sampling
        self.nodes_x, self.nodes_y = mesh.get_nodes()
        self.elements = mesh.get_elements()
        self.space = FEMSpace(mesh)
        self.stiffness_matrix = None
        self.mass_matrix = None
        self.load_vector = None
        self.is_assembled = False
        self.last_assembly_time = 0.0

    def assemble_stiffness_matrix(self, coefficient_func=None):
        """K_ij = sum_e c_e A grad(phi_i).grad(phi_j) (fem_solver.pyx:77-148)."""
        t0 = time.time()
        if coefficient_func is None:
            coefficient_func = lambda x, y: 1.0  # noqa: E731  (reference default)
        self._last_funcs["stiffness"] = coefficient_func
        c = _centroid_or_quadrature(self.space, coefficient_func, self.coefficient_sampling)
        vals = self.space.assemble(c=c)[0]
        self.stiffness_matrix = self.space.to_scipy(vals)
        self.last_assembly_time = time.time() - t0
        return self.stiffness_matrix

    def assemble_mass_matrix(self, density_func=None):
        """M_ij = sum_e rho_e int phi_i phi_j (fem_solver.pyx:150-221), SI area units."""
        t0 = time.time()
        if density_func is None:
            density_func = lambda x, y: 1.0  # noqa: E731
        self._last_funcs["mass"] = density_func
        r = _centroid_or_quadrature(self.space, density_func, self.coefficient_sampling)
        vals = self.space.assemble(r=r, stiffness_mass=False, mass=True)[1]
        self.mass_matrix = self.space.to_scipy(vals, NM * NM)
        self.last_assembly_time += time.time() - t0
        return self.mass_matrix

    def adapt_mesh(self, eigenvector, threshold, cache_dir=""):
        """FEMSolver::adapt_mesh (backend/src/fem.cpp:486-542): flag the elements whose
        least-squares gradient of ``eigenvector`` (nodal values, 1/m with SI
        coordinates) exceeds ``threshold``, refine them (conforming red-green,
        ``adapt.refine_mesh``), rebuild the FE space and reassemble the matrices
        assembled before.  With ``cache_dir`` the refined mesh is saved as
        ``mesh_<|eigenvector|>.txt`` in the reference text format and loaded from
        there on a later call with the same vector norm (fem.cpp:488-500).
        Returns the element flags of the old mesh."""
        import os
        from .adapt import gradient_indicator, refine_mesh
        from .mesh_io import load_mesh, save_mesh
        psi = np.asarray(eigenvector, dtype=np.float64)
        if psi.shape != (self.num_nodes,):
            raise ValueError(f"eigenvector has {psi.size} values, the mesh {self.num_nodes} nodes")
        sp_ = self.space
        flags = (gradient_indicator(sp_.conn, sp_.x, sp_.y, psi, sp_.npe) > threshold).cpu().numpy()
        cache = os.path.join(cache_dir, f"mesh_{np.linalg.norm(psi):.6f}.txt") if cache_dir else ""
        if cache and os.path.exists(cache):
            new = load_mesh(cache)
        else:
            new = refine_mesh(self.mesh, flags)
            if cache:
                os.makedirs(cache_dir, exist_ok=True)
                save_mesh(new, cache)
        funcs = dict(self._last_funcs)
        self.__init__(new, self.element_order, self.coefficient_sampling)
        if "stiffness" in funcs:
            self.assemble_stiffness_matrix(funcs["stiffness"])
        if "mass" in funcs:
            self.assemble_mass_matrix(funcs["mass"])
        return flags

    def assemble_load_vector(self, source_func):
        """f_i = int s phi_i (consistent load; equals fem_solver.pyx:223-270's
        s_e A/3 per vertex for P1 with centroid sampling)."""
        s = _centroid_or_quadrature(self.space, source_func, self.coefficient_sampling)
        vals = self.space.assemble(r=s, stiffness_mass=False, mass=True)[1]
        M = self.space.to_scipy(vals, NM * NM)
        # int s phi_i = sum_j M_s[i, j] * 1 (partition of unity)
        load = np.asarray(M.sum(axis=1)).ravel()
        self.load_vector = load
        return load

    def apply_dirichlet_bc(self, boundary_nodes, boundary_values, matrix, rhs=None):
        """Row replacement as fem_solver.pyx:322-351: each listed row becomes the
        identity row (its other stored entries are removed), columns untouched."""
        A = matrix.tocsr()
        rhs_bc = rhs.copy() if rhs is not None else np.zeros(self.num_nodes)
        nodes = np.asarray(list(boundary_nodes), dtype=np.int64)
        vals = np.asarray(list(boundary_values), dtype=np.float64)
        if nodes.size == 0:
            return A.copy(), rhs_bc
        n = A.shape[0]
        is_bc = np.zeros(n, dtype=bool)
        is_bc[nodes] = True
        lens = np.diff(A.indptr)
        new_lens = np.where(is_bc, 1, lens)
        indptr = np.concatenate([[0], np.cumsum(new_lens)]).astype(A.indptr.dtype)
        indices = np.empty(indptr[-1], dtype=A.indices.dtype)
        data = np.empty(indptr[-1], dtype=np.result_type(A.data.dtype, np.float64))
        keep_rows = ~is_bc
        src_mask = np.repeat(keep_rows, lens)
        dst_mask = np.repeat(keep_rows, new_lens)
        indices[dst_mask] = A.indices[src_mask]
        data[dst_mask] = A.data[src_mask]
        bc_rows = np.nonzero(is_bc)[0]
        indices[indptr[bc_rows]] = bc_rows
        data[indptr[bc_rows]] = 1.0
        rhs_bc[nodes] = vals
        return sp.csr_matrix((data, indices, indptr), shape=A.shape), rhs_bc

    def solve_linear_system(self, matrix, rhs, method='direct'):
        """fem_solver.pyx:353-390; 'cg' runs the HIP CG kernels, 'direct' uses the
        device CG as well when the matrix is symmetric, otherwise GMRES."""
        from .linsolve import solve_linear_system
        if method not in ('direct', 'cg', 'gmres'):
            raise ValueError(f"Unknown solver method: {method}")
        return solve_linear_system(matrix, rhs, method)

    def get_assembly_time(self):
        return self.last_assembly_time

    def get_matrix_info(self):
        info = {
            'num_nodes': self.num_nodes,
            'num_elements': self.num_elements,
            'element_order': self.element_order,
            'assembly_time': self.last_assembly_time,
        }
        if self.stiffness_matrix is not None:
            info['stiffness_nnz'] = self.stiffness_matrix.nnz
        if self.mass_matrix is not None:
            info['mass_nnz'] = self.mass_matrix.nnz
        return info

    def compute_element_quality(self):
        """4 sqrt(3) A / perimeter^2 per element (fem_solver.pyx:392-431)."""
        el = self.space.conn_host[:, :3]
        x, y = self.nodes_x, self.nodes_y
        x0, x1, x2 = x[el[:, 0]], x[el[:, 1]], x[el[:, 2]]
        y0, y1, y2 = y[el[:, 0]], y[el[:, 1]], y[el[:, 2]]
        area = 0.5 * np.abs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
        per = (np.hypot(x1 - x0, y1 - y0) + np.hypot(x2 - x1, y2 - y1) + np.hypot(x0 - x2, y0 - y2))
        q = np.zeros_like(area)
        ok = per > 1e-15
        q[ok] = 4.0 * np.sqrt(3.0) * area[ok] / (per[ok] * per[ok])
        return q


def create_fem_solver(mesh, element_order=1):
    return FEMSolver(mesh, element_order)
