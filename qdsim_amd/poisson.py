"""Poisson equation on the GPU: assembly + Dirichlet lifting + FP64 CG.

Call surface of CythonPoissonSolver (qdsim_cython/solvers/poisson_solver.pyx:31-340):
``PoissonSolver(mesh, epsilon_r_func, rho_func)``, ``solve(V_p, V_n, n=None, p=None)``,
``get_potential()``, public ``stiffness_matrix``/``mass_matrix``.  Discretisation as
in the reference: stiffness with epsilon = eps_r(centroid) EPSILON_0 (:130-140),
consistent mass, Dirichlet V_p on x = 0 and V_n on x = Lx (:276-306), natural
boundary elsewhere.  Right-hand side: M rho_nodal by default (``sign="physical"``,
the weak form of -div(eps grad phi) = rho); ``sign="reference"`` reproduces the
reference's M (-rho_nodal) (:267-274), i.e. K phi = -M rho, under which extra
electrons raise phi and, through V - q phi, attract more electrons.  The reference factorises the row-replaced matrix with spsolve
(:241); here the Dirichlet values are lifted into the right-hand side, the
remaining SPD system is solved with the device CG, and the solution of the
row-replaced system is identical in exact arithmetic.  On a structured mesh the
unknowns are numbered in lattice order and the CG is preconditioned by one
multigrid V-cycle per iteration (qds_cg_mg, PoissonLatticeMG); other meshes use
Jacobi (qds_cg).
"""
import time
import warnings

import numpy as np
import torch

from . import kernels
from .coefficients import evaluate
from .fem import FEMSpace, NM
from .units import EPSILON_0, Q_E


class PoissonProblem:
    """Device-resident Poisson operator of one mesh (values in SI)."""

    def __init__(self, mesh, epsilon_r_func, sampling="centroid", sign="physical", eps_elem=None, contacts=None):
        """eps_elem: per-element permittivity (F per length unit), used instead of
        epsilon_r_func * EPSILON_0; contacts: (left, right) node-id arrays of the V_p
        and V_n contacts instead of the first and last lattice columns."""
        if sign not in ("physical", "reference"):
            raise ValueError(f"sign must be 'physical' or 'reference', got {sign!r}")
        self.sign = sign
        self.mesh = mesh
        self.space = FEMSpace(mesh)
        sp_ = self.space
        if eps_elem is not None:
            eps = torch.as_tensor(np.asarray(eps_elem, dtype=np.float64)).to(sp_.dev).contiguous()
            if eps.shape != (sp_.n_elem,):
                raise ValueError(f"eps_elem needs one value per element ({sp_.n_elem}), got {tuple(eps.shape)}")
        else:
            where = "centroid" if sampling == "centroid" else "quadrature"
            eps = sp_.sample(epsilon_r_func, where, EPSILON_0)
        K, M, _ = sp_.assemble(c=eps, stiffness_mass=True, mass=True)
        self.K = K                       # SI (scale free in 2-D)
        self.M = M * (NM * NM)           # m^2
        conn = mesh.get_elements()
        used = np.zeros(mesh.num_nodes, dtype=bool)
        used[conn.ravel()] = True
        lattice = hasattr(mesh, "lattice_shape") and hasattr(mesh, "lattice_x")
        if lattice:
            ry, rx = mesh.lattice_shape
            lx = np.asarray(mesh.lattice_x)
        if contacts is not None:
            self.left = np.asarray(contacts[0], dtype=np.int64)
            self.right = np.asarray(contacts[1], dtype=np.int64)
        elif lattice:
            self.left = np.nonzero((lx == 0) & used)[0]
            self.right = np.nonzero((lx == rx - 1) & used)[0]
        else:
            # contacts by coordinate: the nodes on x = min x and x = max x
            xn = np.asarray(mesh.get_nodes()[0])
            tol = 1e-9 * (xn[used].max() - xn[used].min())
            self.left = np.nonzero((np.abs(xn - xn[used].min()) <= tol) & used)[0]
            self.right = np.nonzero((np.abs(xn - xn[used].max()) <= tol) & used)[0]
        fixed = np.zeros(mesh.num_nodes, dtype=bool)
        fixed[self.left] = True
        fixed[self.right] = True
        free = np.nonzero(~fixed & used)[0]
        # on a structured mesh the unknowns are lattice columns 1 .. rx-2 of every
        # row: numbered row by row, the system gets a multigrid preconditioner
        self.mg = None
        if lattice and hasattr(mesh, "lattice_y") and rx >= 3:
            ly = np.asarray(mesh.lattice_y)
            key = ly[free] * (rx - 2) + (lx[free] - 1)
            if free.size == (rx - 2) * ry and np.array_equal(np.sort(key), np.arange(free.size)):
                free = free[np.argsort(key)]
                from .lattice import Lattice, PoissonLatticeMG
                eps_e = eps.mean(dim=1) if eps.dim() == 2 else eps
                self.mg = PoissonLatticeMG(Lattice(mesh), eps_e)
        self.free = free
        new = np.full(mesh.num_nodes, -1, dtype=np.int64)
        new[free] = np.arange(free.size)
        r_ip, r_ix, r_src = kernels.pattern_restrict(sp_.indptr_host, sp_.indices_host, new, free.size)
        dev = sp_.dev
        self.r_indptr = torch.from_numpy(r_ip).to(dev)
        self.r_indices = torch.from_numpy(r_ix).to(dev)
        self.r_src = torch.from_numpy(r_src).to(dev)   # restricted CSR slot -> full CSR slot
        self.K_ff = kernels.csr_take(self.r_src, K)
        self.free_d = torch.from_numpy(free).to(dev)
        self.left_d = torch.from_numpy(self.left).to(dev)
        self.right_d = torch.from_numpy(self.right).to(dev)
        self.n = mesh.num_nodes
        self.last_iterations = 0
        self.last_relres = 0.0
        self.last_converged = True

    def spmv(self, vals, x):
        y = torch.empty_like(x)
        kernels.spmm(self.space.indptr, self.space.indices, self.n, self.space.nnz, vals, x, y)
        return y

    def solve(self, rho_nodal, V_p, V_n, x0=None, rtol=1e-12):
        """phi (device float64, SI volts) for nodal charge density rho (C/m^3)."""
        rho = rho_nodal.contiguous()
        rhs = self.spmv(self.M, rho if self.sign == "physical" else -rho)
        g = torch.zeros(self.n, dtype=torch.float64, device=rhs.device)
        g[self.left_d] = V_p
        g[self.right_d] = V_n
        t = self.spmv(self.K, g)
        b = (rhs - t)[self.free_d].contiguous()
        x = self.solve_free(b, None if x0 is None else x0[self.free_d], rtol)
        phi = g.clone()
        phi[self.free_d] = x
        return phi

    def solve_free(self, b, x0=None, rtol=1e-12):
        """x with K_ff x = b on the free DOFs (ordering of ``free``): device CG,
        multigrid-preconditioned on structured meshes.  K_ff is symmetric, so the
        adjoint of the Poisson step solves with the same call."""
        b = b.contiguous()
        x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
        if self.mg is not None:
            it, rr = kernels.cg_mg(self.r_indptr, self.r_indices, self.K_ff, self.free.size,
                                   int(self.r_indices.numel()), b, x, self.mg._structs, self.mg.n_levels,
                                   self.mg.nu, self.mg.coarse_sweeps, self.mg.omega, rtol=rtol, max_iter=2000)
        else:
            it, rr = kernels.cg(self.r_indptr, self.r_indices, self.K_ff, self.free.size,
                                int(self.r_indices.numel()), b, x, rtol=rtol, max_iter=20000)
        self.last_iterations, self.last_relres = it, rr
        self.last_converged = rr <= rtol
        if not self.last_converged:
            # spsolve (poisson_solver.pyx:241) is direct; an iterative solve that
            # stops short is reported rather than passed on silently
            warnings.warn(f"Poisson CG stopped after {it} iterations at relative residual {rr:.3e} "
                          f"(rtol {rtol:.1e})", RuntimeWarning, stacklevel=2)
        return x


class PoissonSolver:
    """Drop-in for CythonPoissonSolver (poisson_solver.pyx:31)."""

    def __init__(self, mesh, epsilon_r_func, rho_func, sign="physical"):
        self.num_nodes = mesh.num_nodes
        self.num_elements = mesh.num_elements
        self.Lx, self.Ly = mesh.Lx, mesh.Ly
        self.nx, self.ny = mesh.nx, mesh.ny
        self.mesh = mesh
        self.nodes_x, self.nodes_y = mesh.get_nodes()
        self.epsilon_r_func = epsilon_r_func
        self.rho_func = rho_func
        self.problem = PoissonProblem(mesh, epsilon_r_func, sign=sign)
        self.stiffness_matrix = self.problem.space.to_scipy(self.problem.K)
        self.mass_matrix = self.problem.space.to_scipy(self.problem.M)
        self.potential = np.zeros(self.num_nodes)
        self.electric_field_x = np.zeros(self.num_nodes)
        self.electric_field_y = np.zeros(self.num_nodes)
        self.is_assembled = True
        self.last_solve_time = 0.0

    def _rho_nodal(self, n, p):
        if n is None:
            n = np.zeros(self.num_nodes)
        if p is None:
            p = np.zeros(self.num_nodes)
        try:
            v = np.asarray(self.rho_func(self.nodes_x, self.nodes_y, n, p), dtype=np.float64)
            if v.shape == (self.num_nodes,):
                return v
        except Exception:
            pass
        return np.array([self.rho_func(float(x), float(y), n, p)
                         for x, y in zip(self.nodes_x, self.nodes_y)], dtype=np.float64)

    def solve(self, V_p, V_n, n=None, p=None):
        t0 = time.time()
        rho = torch.from_numpy(self._rho_nodal(n, p)).to(self.problem.space.dev)
        phi = self.problem.solve(rho, float(V_p), float(V_n))
        self.potential = phi.cpu().numpy()
        self._compute_electric_field()
        self.last_solve_time = time.time() - t0

    def _compute_electric_field(self):
        """Finite differences on the vertex grid (poisson_solver.pyx:308-335); P1 only."""
        if getattr(self.mesh, "element_order", 1) != 1:
            return
        dx = self.Lx / (self.nx - 1)
        dy = self.Ly / (self.ny - 1)
        P = self.potential.reshape(self.ny, self.nx)
        Ex = np.empty_like(P)
        Ey = np.empty_like(P)
        Ex[:, 1:-1] = -(P[:, 2:] - P[:, :-2]) / (2 * dx)
        Ex[:, 0] = -(P[:, 1] - P[:, 0]) / dx
        Ex[:, -1] = -(P[:, -1] - P[:, -2]) / dx
        Ey[1:-1] = -(P[2:] - P[:-2]) / (2 * dy)
        Ey[0] = -(P[1] - P[0]) / dy
        Ey[-1] = -(P[-1] - P[-2]) / dy
        self.electric_field_x = Ex.ravel()
        self.electric_field_y = Ey.ravel()

    def get_potential(self):
        return self.potential.copy()

    def get_electric_field(self):
        return self.electric_field_x.copy(), self.electric_field_y.copy()
