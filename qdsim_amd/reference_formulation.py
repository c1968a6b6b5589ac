"""FixedOpenSystemSolver(formulation='reference'): the reference's own formulas.

Element matrices of qdsim_cython/solvers/fixed_open_system_solver.pyx:197-271
(centroid m*, V; V += CAP absorption when open; kinetic hbar^2/(2m) A/3 on the
diagonal and half of it off the diagonal; potential V A/3 on the diagonal;
M = A/12 (1 + delta_ij)), assembled on the GPU through the same segmented gather
as the FEM path.  Boundary treatment of :273-291: the rows of nodes 0 and N-1
become identity rows in H and M.  For any eigenvalue other than 1 (J) such a
pencil has psi_0 = psi_{N-1} = 0, and its remaining rows are exactly the
symmetric pencil restricted to the other nodes, so that restricted pencil is
solved; ARPACK's which='SM' (:302-305) is honoured by taking the eigenvalues of
smallest magnitude.  Open systems get the constant shift E - i 0.1 cap_strength
of :321-328.
"""
import numpy as np
import torch

from . import kernels
from .eigen import gpu_eigsh
from .units import EV_TO_J, HBAR


def reference_element_matrices(solver, space):
    """Packed (E, 6) H and M element matrices in SI (J, m^2)."""
    xc, yc = space.centroids()                       # nm
    x_m, y_m = xc * 1e-9, yc * 1e-9
    from .coefficients import evaluate
    m_star = evaluate(solver.m_star_func, x_m, y_m)
    V = evaluate(solver.potential_func, x_m, y_m)
    if solver.use_open_boundaries:
        V = V + solver.cap_absorption(x_m, y_m)
    conn = space.conn.long()
    X = space.x[conn[:, :3]] * 1e-9
    Y = space.y[conn[:, :3]] * 1e-9
    area = 0.5 * torch.abs((X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (Y[:, 1] - Y[:, 0]))
    kin = HBAR * HBAR / (2.0 * m_star) * area / 3.0
    pot = V * area / 3.0
    mass = area / 12.0
    # packed upper triangle order: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
    diag_pos = [0, 3, 5]
    off_pos = [1, 2, 4]
    He = torch.empty((area.shape[0], 6), dtype=torch.float64, device=area.device)
    Me = torch.empty_like(He)
    for p in diag_pos:
        He[:, p] = kin + pot
        Me[:, p] = 2.0 * mass
    for p in off_pos:
        He[:, p] = 0.5 * kin
        Me[:, p] = mass
    valid = (area >= 1e-20).to(torch.float64)[:, None]   # :209-210 skips degenerate elements
    return (He * valid).reshape(-1).contiguous(), (Me * valid).reshape(-1).contiguous()


def reference_matrices(solver):
    """Full assembled (H, M) scipy CSR in SI, before boundary rows."""
    from .fem import FEMSpace
    if solver.element_order != 1:
        raise ValueError("the reference formulation is P1 only")
    space = FEMSpace(solver.mesh)
    He, Me = reference_element_matrices(solver, space)
    H, M, _ = kernels.assemble_gather(space.nnz, space.seg_ptr, space.seg_src, He, Me)
    return space, space.to_scipy(H), space.to_scipy(M)


def solve_reference_formulation(solver, k):
    space, H, M = reference_matrices(solver)
    n = H.shape[0]
    keep = np.ones(n, dtype=bool)
    keep[[0, n - 1]] = False
    free = np.nonzero(keep)[0]
    Hf = H[free][:, free].tocsr()
    Mf = M[free][:, free].tocsr()
    k_req = max(1, min(k, n - 3))                    # :298-300
    extra = min(len(free) // 3 - k_req, max(4, k_req))
    vals, vecs, _ = gpu_eigsh(Hf, Mf, k_req + max(extra, 0), which="SA", tol=1e-10)
    order = np.argsort(np.abs(vals))[:k_req]         # which='SM'
    order = order[np.argsort(vals[order])]
    vals = vals[order]
    full = np.zeros((n, k_req))
    full[free] = vecs[:, order]
    if solver.use_open_boundaries:
        vals = vals - 1j * solver.cap_strength * 0.1
    return vals, full
