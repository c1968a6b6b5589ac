"""FP64 Poisson oracle (test use only).

Restates qdsim_cython/solvers/poisson_solver.pyx:94-306: P1 stiffness with
epsilon = eps_r(centroid) * EPSILON_0, consistent mass A/12 (1 + delta), RHS
M @ (-rho_nodal) (sign as in the reference, :267-274), Dirichlet V_p at x = 0 and
V_n at x = Lx by row replacement, spsolve.  Pinned by the exact linear solution
phi = V_p + (V_n - V_p) x / Lx for rho = 0 (poisson_solver.pyx:393-403 test case)
and, for uniform rho, by the nodally exact quadratic
phi = V_p + (V_n - V_p) x / L -/+ rho / (2 eps) x (x - L) (reference / physical
sign), which fixes the sign and the mass weighting of the charge term.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import fem_ref, mesh_ref

EPSILON_0 = 8.854187817e-12


def poisson_matrices(nx, ny, Lx, Ly, eps_r_func):
    x, y, el = mesh_ref.simple_mesh_p1(nx, ny, Lx, Ly)
    vx, vy = x[el], y[el]
    xc, yc = vx.mean(1), vy.mean(1)
    eps = eps_r_func(xc, yc) * EPSILON_0 * np.ones_like(xc)
    K, Mm = fem_ref.p1_centroid_matrices(vx, vy, eps, np.ones_like(xc))
    n = len(x)
    return x, y, el, fem_ref.assemble(el, n, K), fem_ref.assemble(el, n, Mm)


def solve(nx, ny, Lx, Ly, eps_r_func, rho_nodal, V_p, V_n, sign="reference"):
    """sign "reference": RHS M (-rho) as poisson_solver.pyx:267-274, i.e. K phi =
    -M rho; "physical": RHS M rho, the weak form of -div(eps grad phi) = rho."""
    x, y, el, K, M = poisson_matrices(nx, ny, Lx, Ly, eps_r_func)
    if sign not in ("reference", "physical"):
        raise ValueError(sign)
    rhs = M @ (-rho_nodal if sign == "reference" else rho_nodal)
    A = K.tolil()
    b = rhs.copy()
    for i in range(len(x)):
        if abs(x[i]) < 1e-12:
            A[i, :] = 0; A[i, i] = 1; b[i] = V_p
        elif abs(x[i] - Lx) < 1e-12:
            A[i, :] = 0; A[i, i] = 1; b[i] = V_n
    return spla.spsolve(A.tocsr(), b), x, y
