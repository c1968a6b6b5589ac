"""Device linear solves behind FEMSolver.solve_linear_system.

The reference calls scipy spsolve ('direct'), or cg / gmres with the removed
``tol=`` keyword (fem_solver.pyx:353-390, which raises TypeError with scipy >= 1.14).
Here every method runs the FP64 device CG: rows that are identity rows (the
Dirichlet rows written by apply_dirichlet_bc) are lifted into the right-hand
side, and the remaining block must be symmetric positive definite (stiffness
matrices, shifted operators); a non-symmetric remainder raises ValueError.
"""
import numpy as np
import scipy.sparse as sp
import torch

from . import kernels
from .device import device


def _identity_rows(A):
    lens = np.diff(A.indptr)
    rows = np.nonzero(lens == 1)[0]
    if rows.size == 0:
        return rows
    pos = A.indptr[rows]
    ok = (A.indices[pos] == rows) & (A.data[pos] == 1.0)
    return rows[ok]


def solve_linear_system(matrix, rhs, method="direct", rtol=1e-12):
    A = sp.csr_matrix(matrix, dtype=np.float64)
    A.sort_indices()
    b = np.asarray(rhs, dtype=np.float64)
    n = A.shape[0]
    fixed = _identity_rows(A)
    x = np.zeros(n)
    x[fixed] = b[fixed]
    free_mask = np.ones(n, dtype=bool)
    free_mask[fixed] = False
    free = np.nonzero(free_mask)[0]
    if free.size == 0:
        return x
    Aff = A[free][:, free].tocsr()
    Aff.sort_indices()
    asym = abs(Aff - Aff.T).max() if Aff.nnz else 0.0
    if asym > 1e-12 * (abs(Aff).max() + 1e-300):
        raise ValueError("device CG needs a symmetric system after Dirichlet lifting")
    rhs_f = b[free] - A[free][:, fixed] @ x[fixed]
    dev = device()
    ip = torch.from_numpy(Aff.indptr.astype(np.int32)).to(dev)
    ix = torch.from_numpy(Aff.indices.astype(np.int32)).to(dev)
    v = torch.from_numpy(Aff.data).to(dev)
    bt = torch.from_numpy(rhs_f).to(dev)
    xt = torch.zeros_like(bt)
    it, rr = kernels.cg(ip, ix, v, free.size, Aff.nnz, bt, xt, rtol=rtol, max_iter=50000)
    if rr > 1e3 * rtol:
        print(f"CG solver convergence issue (relative residual {rr:.2e} after {it} iterations)")
    x[free] = xt.cpu().numpy()
    return x
