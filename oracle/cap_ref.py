"""Complex-symmetric CAP oracle (test use only).

Dense LAPACK generalized eigensolve (scipy.linalg.eig -> zggev) of the restricted
pencil (H - i W, M) built by the FEM oracle; all eigenvalues, so the selection by
real part or by distance to a target is exact.  Pinned by: W = 0 reduces to the
real symmetric problem (eig_ref), the pencil is complex symmetric, and every
eigenvalue has Im E <= 0 for W >= 0 (absorption)."""
import numpy as np
import scipy.linalg as sla

from . import eig_ref


def cap_dense(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, cap_func):
    m = eig_ref.schrodinger_matrices(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, cap_func=cap_func)
    free = eig_ref.free_dofs(m["x"], m["y"], m["el"], Lx_nm, Ly_nm)
    H = m["H"][free][:, free].toarray()
    W = m["W"][free][:, free].toarray()
    M = m["M"][free][:, free].toarray()
    A = H - 1j * W
    vals, vecs = sla.eig(A, M)
    order_ = np.argsort(vals.real)
    m.update(vals=vals[order_], vecs_free=vecs[:, order_], free=free, A=A, Md=M)
    return m


def cap_eigs(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, cap_func, k, sigma):
    """The k eigenpairs of (H - i W, M) nearest sigma (eV) by ARPACK shift-invert with
    a sparse LU, scipy.sparse.linalg.eigs(A, k, M, sigma): the CPU call the CAP
    configuration names.  Sorted by distance to sigma.  Pinned against cap_dense on
    small meshes (tests/test_oracle_pinning.py)."""
    import scipy.sparse.linalg as spla
    m = eig_ref.schrodinger_matrices(nx, ny, Lx_nm, Ly_nm, order, mstar_func, v_func, cap_func=cap_func)
    free = eig_ref.free_dofs(m["x"], m["y"], m["el"], Lx_nm, Ly_nm)
    A = (m["H"][free][:, free] - 1j * m["W"][free][:, free]).tocsc()
    M = m["M"][free][:, free].tocsc().astype(np.complex128)
    vals, vecs = spla.eigs(A, k=k, M=M, sigma=sigma, which="LM", tol=1e-12)
    order_ = np.argsort(np.abs(vals - sigma))
    m.update(vals=vals[order_], vecs_free=vecs[:, order_], free=free, A=A, Ms=M)
    return m
