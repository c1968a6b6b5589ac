"""FP64 eigen oracle (test use only).

Reference call sites: scipy eigsh in eigen_solver.pyx:225-235 (which='SM', tol 1e-8)
and the large-N recipe eigsh(H, k, M, sigma, which='LM') of
qdsim_cython/advanced_eigenvalue_solvers.py:106-122.  The oracle uses the
shift-invert form with a tight tolerance.  Pinned by the tridiagonal spectrum of
eigen_solver.pyx:443-447 and the particle-in-a-box formula of
eigen_solver.pyx:350-354 (tests/test_oracle_pinning.py).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import fem_ref, mesh_ref

HBAR = 1.054571817e-34
M_E = 9.1093837015e-31
EV = 1.602176634e-19
KIN = HBAR * HBAR / 2.0 / EV / 1e-18      # hbar^2/2 [eV nm^2 kg]


def lowest(H, M, k, sigma, tol=1e-13):
    vals, vecs = spla.eigsh(H.tocsc(), k=k, M=M.tocsc(), sigma=sigma, which='LM', tol=tol)
    order = np.argsort(vals)
    return vals[order], vecs[:, order]


def reference_eigsh_sm(H, M, k, tol=1e-8, maxiter=1000):
    """The reference's exact call (eigen_solver.pyx:225-235)."""
    return spla.eigsh(H, k=k, M=M, which='SM', tol=tol, maxiter=maxiter, sigma=None)


def schrodinger_matrices(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, sampling="quadrature",
                         cap_func=None):
    """Oracle H (eV), M (nm^2) [and W] on the reference mesh, full DOF numbering.

    mstar_func(x_nm, y_nm) -> m*/m0 ; v_func(x_nm, y_nm) -> eV (numpy vectorised)."""
    x, y, el = mesh_ref.reference_mesh(nx, ny, Lx_nm, Ly_nm, order)
    vx, vy = x[el[:, :3]], y[el[:, :3]]
    deg = 2 * order + 2
    if sampling == "quadrature":
        xq, yq = fem_ref.quad_points(vx, vy, deg)
    else:
        xq = vx.mean(1, keepdims=True) * np.ones((1, len(fem_ref.rule(deg)[1])))
        yq = vy.mean(1, keepdims=True) * np.ones((1, len(fem_ref.rule(deg)[1])))
    c_q = KIN / (mstar_func(xq, yq) * M_E)
    v_q = v_func(xq, yq) * np.ones_like(xq)
    K, Vm, Mm = fem_ref.element_matrices(order, vx, vy, deg, c_q=c_q, v_q=v_q)
    n = len(x)
    H = fem_ref.assemble(el, n, K + Vm)
    M = fem_ref.assemble(el, n, Mm)
    out = {"x": x, "y": y, "el": el, "H": H, "M": M}
    if cap_func is not None:
        w_q = cap_func(xq, yq) * np.ones_like(xq)
        _, Wm, _ = fem_ref.element_matrices(order, vx, vy, deg, c_q=None, v_q=w_q)
        out["W"] = fem_ref.assemble(el, n, Wm)
    return out


def free_dofs(x, y, el, Lx, Ly):
    used = np.zeros(len(x), bool)
    used[el.ravel()] = True
    tol = 1e-9 * max(Lx, Ly)
    bnd = (np.abs(x) < tol) | (np.abs(x - Lx) < tol) | (np.abs(y) < tol) | (np.abs(y - Ly) < tol)
    return np.nonzero(used & ~bnd)[0]


def solve_closed(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, k, sampling="quadrature"):
    m = schrodinger_matrices(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, sampling)
    free = free_dofs(m["x"], m["y"], m["el"], Lx_nm, Ly_nm)
    Hf = m["H"][free][:, free]
    Mf = m["M"][free][:, free]
    vmin = float(v_func(m["x"], m["y"]).min()) if np.ndim(v_func(m["x"], m["y"])) else float(v_func(0, 0))
    vals, vecs = lowest(Hf, Mf, k, sigma=vmin - 0.05)
    full = np.zeros((len(m["x"]), k))
    full[free] = vecs
    m.update(vals=vals, vecs=full, free=free)
    return m
