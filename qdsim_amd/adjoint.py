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

Degenerate levels: within a cluster of (near-)equal eigenvalues the individual
E_k are not differentiable and the per-state weights psi_k^2 depend on the
basis the solver returned; only symmetric functions of the cluster (e.g. its
sum, or any loss that weighs its members equally) have well-defined gradients,
sum_k psi_k^2 being basis independent.  ``degenerate_clusters`` finds the
clusters, and the backward warns when the incoming gradient differs inside one.

Gate and doping parameters act through the Poisson step, which is linear:
phi_f = K_ff^-1 P_f (s M rho - K g), phi_d = g_d (g: the Dirichlet values of the
contacts, s = +1 for the physical sign, -1 for the reference's).  For an
objective J(phi) one adjoint solve K_ff lambda = dJ/dphi_f (K symmetric, the same
device CG as the forward solve) gives
    dJ/drho = s M lambda_ext,   dJ/dg_d = dJ/dphi_d - (K lambda_ext)_d
(lambda_ext: lambda on the free DOFs, 0 on the contacts).  ``poisson_potential``
is that step as a torch.autograd.Function of (rho, V_p, V_n), and
``electrostatic_levels`` chains it with the FE interpolation of phi to the
quadrature points and the Hellmann-Feynman eigenvalues: the levels of V - q phi
as a differentiable function of any torch expression for the charge density
(doping profile) and the contact (gate) voltages.
"""
import warnings

import numpy as np
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


def degenerate_clusters(lam, rel_gap=1e-6):
    """Groups of consecutive (sorted) eigenvalues closer than rel_gap times the
    spread of lam (or of |lam| for a single value): lists of indices, only groups
    of two or more."""
    lam = np.asarray(lam, dtype=np.float64)
    if lam.size < 2:
        return []
    scale = max(float(lam.max() - lam.min()), float(np.abs(lam).max()), 1e-300)
    out, cur = [], [0]
    for i in range(1, lam.size):
        if lam[i] - lam[i - 1] <= rel_gap * scale:
            cur.append(i)
        else:
            if len(cur) > 1:
                out.append(cur)
            cur = [i]
    if len(cur) > 1:
        out.append(cur)
    return out


class _Eigenvalues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v_q, prob, k, mode, tol):
        prob.set_potential_ev(v_q.detach().contiguous())
        prob.assemble()
        lam, psi, info = prob.solve(k, mode=mode, tol=tol)
        rho = state_density_q(prob.space, psi)
        ctx.save_for_backward(rho)
        ctx.clusters = degenerate_clusters(lam)
        prob.last_eigenvectors = psi
        prob.last_info = info
        return torch.as_tensor(lam, dtype=torch.float64, device=v_q.device)

    @staticmethod
    def backward(ctx, gE):
        (rho,) = ctx.saved_tensors
        if ctx.clusters:
            g = gE.detach().double().cpu().numpy()
            for c in ctx.clusters:
                if np.ptp(g[c]) > 1e-12 * max(np.abs(g).max(), 1e-300):
                    warnings.warn(f"gradient through degenerate levels {c}: the incoming gradient differs "
                                  f"inside the cluster, so the result depends on the eigenvector basis; "
                                  f"use a loss symmetric in these levels", RuntimeWarning, stacklevel=2)
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
    return poisson_problem.solve_free(dJ_dphi_free, rtol=rtol)


class _PoissonPotential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, V_p, V_n, pp, rtol):
        ctx.pp, ctx.rtol = pp, rtol
        return pp.solve(rho.detach().contiguous(), float(V_p), float(V_n), rtol=rtol)

    @staticmethod
    def backward(ctx, g_phi):
        pp = ctx.pp
        g_phi = g_phi.contiguous()
        lam = torch.zeros_like(g_phi)
        lam[pp.free_d] = poisson_adjoint(pp, g_phi[pp.free_d], ctx.rtol)
        s = 1.0 if pp.sign == "physical" else -1.0
        g_rho = s * pp.spmv(pp.M, lam)
        g_dir = g_phi - pp.spmv(pp.K, lam)
        g_vp = g_dir[pp.left_d].sum().reshape(())
        g_vn = g_dir[pp.right_d].sum().reshape(())
        return g_rho, g_vp, g_vn, None, None


def poisson_potential(pp, rho_nodal, V_p, V_n, rtol=1e-12):
    """Nodal phi (V) of the Poisson problem ``pp`` (a PoissonProblem) for the charge
    density rho_nodal (C/m^3) and contact voltages V_p (x = 0), V_n (x = Lx),
    differentiable in all three (tensors; plain numbers are taken as constants)."""
    dev = rho_nodal.device
    V_p = torch.as_tensor(V_p, dtype=torch.float64, device=dev)
    V_n = torch.as_tensor(V_n, dtype=torch.float64, device=dev)
    return _PoissonPotential.apply(rho_nodal, V_p, V_n, pp, rtol)


def electrostatic_levels(prob, pp, v0_q, rho_nodal, V_p, V_n, k, mode="mixed", tol=1e-9, rtol=1e-12):
    """Lowest k levels (eV) of V0 - q phi, phi from the Poisson problem ``pp`` on the
    same mesh as the SchrodingerProblem ``prob`` (v0_q: the fixed potential at its
    quadrature points, eV).  Differentiable in rho_nodal, V_p and V_n (and v0_q)."""
    if pp.space.n != prob.space.n or pp.space.npe != prob.space.npe:
        raise ValueError("Poisson and Schroedinger problems must share the mesh and element order")
    phi = poisson_potential(pp, rho_nodal, V_p, V_n, rtol)
    return eigenvalues(prob, v0_q - nodal_to_quadrature(prob.space, phi), k, mode, tol)
