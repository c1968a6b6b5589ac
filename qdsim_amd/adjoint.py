"""Differentiable eigenvalues: Hellmann-Feynman adjoint behind torch.autograd.

No reference counterpart (the reference has no gradients).  For the closed
pencil H(V) psi = E M psi with M-normalised real psi,
    dE_k / dV(x_q) = w_q |J_e| / 2 * psi_k(x_q)^2,
the quadrature-point weights produced by the qds_state_density_q kernel.  The
potential enters the assembly only through its values at the quadrature points,
so ``eigenvalues(v_q)`` is an autograd function of the (E, Q) tensor v_q (eV);
any torch expression that produces v_q from gate or doping parameters gets exact
gradients by the chain rule, and ``nodal_to_quadrature`` provides dE/dV_nodal
for a potential given at the DOFs.

The Poisson part of the self-consistent loop is linear: phi = K^-1 (f(theta) -
K g), so d/dtheta of an objective through phi needs one adjoint solve
K lambda = dJ/dphi with the same device CG (``poisson_adjoint``).
"""
import torch

from . import _lib
from .device import ptr, stream_ptr


def state_density_q(space, psi):
    """(E, Q, k) Hellmann-Feynman weights for full-numbering eigenvectors psi (n, k)."""
    psi = psi.contiguous()
    k = psi.shape[1]
    out = torch.empty((space.n_elem, space.nq, k), dtype=torch.float64, device=psi.device)
    _lib.check(_lib.lib.qds_state_density_q(space.npe, space.nq, space.n_elem, ptr(space.conn), ptr(space.x),
                                            ptr(space.y), ptr(space.tables), ptr(psi), psi.stride(0), k,
                                            ptr(out), stream_ptr()), "qds_state_density_q")
    return out


class _Eigenvalues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v_q, prob, k, mode, tol):
        prob.set_potential_ev(v_q.detach().contiguous())
        prob.assemble()
        lam, psi, info = prob.solve(k, mode=mode, tol=tol)
        rho = state_density_q(prob.space, psi)
        ctx.save_for_backward(rho)
        prob.last_eigenvectors = psi
        prob.last_info = info
        return torch.as_tensor(lam, dtype=torch.float64, device=v_q.device)

    @staticmethod
    def backward(ctx, gE):
        (rho,) = ctx.saved_tensors
        gV = torch.einsum("eqk,k->eq", rho, gE.to(rho.dtype))
        return gV, None, None, None, None


def eigenvalues(prob, v_q, k, mode="mixed", tol=1e-9):
    """Lowest k eigenvalues (eV) of the SchrodingerProblem with potential v_q (E, Q) in eV."""
    return _Eigenvalues.apply(v_q, prob, k, mode, tol)


def nodal_to_quadrature(space, v_nodal):
    """V at the quadrature points from DOF values (differentiable torch expression)."""
    phi = space.tables[space.nq: space.nq + space.nq * space.npe].reshape(space.nq, space.npe)
    return v_nodal[space.conn.long()] @ phi.T


def quadrature_points_nm(space):
    xq, yq = space.quad_points()
    return xq, yq


def poisson_adjoint(poisson_problem, dJ_dphi_free, rtol=1e-12):
    """lambda = K_ff^-1 dJ/dphi_free (K symmetric), the adjoint of the Poisson step."""
    from . import kernels
    pp = poisson_problem
    lam = torch.zeros_like(dJ_dphi_free)
    kernels.cg(pp.r_indptr, pp.r_indices, pp.K_ff, pp.free.size, int(pp.r_indices.numel()),
               dJ_dphi_free.contiguous(), lam, rtol=rtol, max_iter=20000)
    return lam
