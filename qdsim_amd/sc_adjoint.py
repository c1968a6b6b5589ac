"""Gradients through the self-consistent Poisson-Schroedinger fixed point.

No reference counterpart (the reference computes no gradients); the map is the
outer iteration of qdsim_cython/solvers/self_consistent_solver.pyx:165-301 as
``SelfConsistentSolver`` runs it:

    V = v0 - phi (eV at the quadrature points),  (E_s, psi_s) lowest k of H(V),
    E_F = E_0 + kT/2,  occ_s = f((E_s - E_F)/kT),  n_raw = sum_s occ_s psi_s^2,
    n = mean(N_D) n_raw / mean(n_raw),  p = N_A exp(-q phi/kT) (phi > 0, else N_A),
    F(phi) = Poisson(q (p - n); V_p, V_n).

At a fixed point phi* = F(phi*) the mixing drops out, and for an objective
J = g_phi . phi* + g_E . E*(phi*) the adjoint of the fixed point is
    mu = (I - dF/dphi)^-T (g_phi + (dE/dphi)^T g_E),
solved by GMRES on the device; each product with (dF/dphi)^T is one adjoint
Poisson solve (the forward CG: K_ff is symmetric) followed by the adjoint of the
density response of the eigenstates:

* occupations: d occ_s = f'_s (dE_s - dE_0) / kT with dE_s = <psi_s, dH psi_s>
  (Hellmann-Feynman weights, qds_state_density_q);
* rotations inside the computed states: the pairs (s, t <= k) enter as
  (occ_s - occ_t)/(E_s - E_t) <psi_s, dH psi_t> psi_s psi_t, finite through
  degenerate pairs (the divided difference becomes f'/kT); the symmetric pair
  matrix is diagonalised so that its weights come from k state densities of
  rotated vectors;
* the rest of the spectrum (t > k) without computing it: for each occupied s one
  projected linear solve (Sternheimer) (H - E_s M) y_s = (I - M Psi Psi^T) r_s with
  y_s M-orthogonal to the computed states, where r_s = 2 occ_s psi_s (.) g; the
  operator is positive definite there (E_t > E_s for every t > k), and the solves
  run as one block CG with a shift per column, preconditioned by the lattice
  multigrid for (H - sigma M), sigma below E_0 (preconditioned spectrum
  (E_t - E_s)/(E_t - sigma) in (0, 1]).  Their contribution to dJ/dV(x_q) is
  -w_q sum_s y_s(x_q) psi_s(x_q), taken from state densities by polarisation.

The parameters: v0 at the quadrature points (any gate potential by the chain
rule), the acceptor and donor profiles N_A, N_D (nodal) and the contact voltages
V_p, V_n.  Requires E_k < E_{k+1} for the states that carry occupation (the
density of k states is otherwise not a differentiable function of V).
"""
import warnings

import numpy as np
import torch

from . import _lib, kernels
from .adjoint import state_density_q
from .device import ptr, stream_ptr, sync
from .units import EV_TO_J, K_B, Q_E

C0 = 1e9          # SI scale of the eigenvectors (M_nm-normalised -> M_SI-normalised)


def _occupations(lam_ev, temperature):
    kT = K_B * temperature
    E = np.asarray(lam_ev, dtype=np.float64) * EV_TO_J
    arg = (E - (E[0] + 0.5 * kT)) / kT
    occ = np.where(arg < 50, 1.0 / (1.0 + np.exp(np.minimum(arg, 50))), 0.0)
    docc = np.where(arg < 50, -occ * (1.0 - occ), 0.0) * (EV_TO_J / kT)     # per eV
    return occ, docc


def _pair_weights(lam, occ, docc):
    """D_st = (occ_s - occ_t) / (E_s - E_t) for s != t (f'/kT through degenerate
    pairs), 0 on the diagonal."""
    k = len(lam)
    D = np.zeros((k, k))
    for s in range(k):
        for t in range(k):
            if s == t:
                continue
            dE = lam[s] - lam[t]
            if abs(dE) > 1e-9 * max(1.0, abs(lam[s])):
                D[s, t] = (occ[s] - occ[t]) / dE
            else:
                D[s, t] = 0.5 * (docc[s] + docc[t])
    return D


def _gmres(matvec, b, rtol, restart=40, maxiter=400):
    """x with matvec(x) = b (device FP64 vectors), restarted GMRES, modified
    Gram-Schmidt; returns (x, iterations, relative residual)."""
    n = b.numel()
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.vector_norm(b))
    if bnorm == 0.0:
        return x, 0, 0.0
    its = 0
    r = b.clone()
    beta = bnorm
    while its < maxiter:
        V = torch.empty((restart + 1, n), dtype=torch.float64, device=b.device)
        V[0] = r / beta
        H = np.zeros((restart + 1, restart))
        g = np.zeros(restart + 1)
        g[0] = beta
        cs, sn = np.zeros(restart), np.zeros(restart)
        j_end = 0
        for j in range(restart):
            its += 1
            w = matvec(V[j])
            for i in range(j + 1):
                H[i, j] = float(torch.dot(w, V[i]))
                w -= H[i, j] * V[i]
            H[j + 1, j] = float(torch.linalg.vector_norm(w))
            if H[j + 1, j] > 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / d, H[j + 1, j] / d
            H[j, j], H[j + 1, j] = d, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            j_end = j + 1
            if abs(g[j + 1]) <= rtol * bnorm or its >= maxiter:
                break
        y = np.linalg.solve(np.triu(H[:j_end, :j_end]), g[:j_end])
        x += torch.from_numpy(y).to(b.device) @ V[:j_end]
        r = b - matvec(x)
        beta = float(torch.linalg.vector_norm(r))
        if beta <= rtol * bnorm:
            break
    return x, its, beta / bnorm


class SCLinearization:
    """The self-consistent map of ``sc`` (a SelfConsistentSolver) linearised at phi
    (nodal, V): eigenpairs recomputed at phi, products with (dF/dphi)^T, and the
    gradients of the fixed point.

    phi: the converged potential (device tensor or array over the mesh nodes);
    V_p, V_n: contact voltages; N_A, N_D: scalars or nodal arrays; k: the number
    of states of the density (``num_quantum_states`` of the solve)."""

    def __init__(self, sc, phi, V_p, V_n, N_A, N_D, k, eig_tol=1e-11, sternheimer_rtol=1e-11,
                 occ_floor=1e-14):
        self.sc = sc
        sp, pp = sc.schrodinger, sc.poisson
        self.sp, self.pp = sp, pp
        dev = sp.space.dev
        self.dev = dev
        self.n = sp.space.n
        self.V_p, self.V_n = float(V_p), float(V_n)
        self.phi = torch.as_tensor(np.asarray(phi.cpu() if torch.is_tensor(phi) else phi, dtype=np.float64),
                                   device=dev).contiguous()
        self.N_A = torch.as_tensor(np.broadcast_to(np.asarray(N_A, dtype=np.float64), (self.n,)).copy(), device=dev)
        self.N_D = torch.as_tensor(np.broadcast_to(np.asarray(N_D, dtype=np.float64), (self.n,)).copy(), device=dev)
        self.k = int(k)
        self.kT = K_B * sc.temperature
        self.rtol_st = float(sternheimer_rtol)
        # eigenpairs at phi
        sp.set_potential_ev((sc.v0_q - sc._phi_at_quadrature(self.phi)).contiguous())
        sp.assemble()
        lam, psi, info = sp.solve(self.k, tol=eig_tol)
        if not info["converged"]:
            raise RuntimeError(f"eigensolve at the fixed point did not converge (residuals {info['residuals']})")
        self.lam = np.asarray(lam, dtype=np.float64)
        self.psi = psi.contiguous()                           # (n, k), M_nm-normalised
        self.occ, self.docc = _occupations(self.lam, sc.temperature)
        self.hf = state_density_q(sp.space, self.psi)        # (E, Q, k): dE_s / dV(x_q)
        # densities at phi
        occ_d = torch.from_numpy(self.occ).to(dev)
        psi_si = (self.psi * C0).contiguous()
        n_raw = torch.empty(self.n, dtype=torch.float64, device=dev)
        _lib.check(_lib.lib.qds_density(self.n, self.k, ptr(psi_si), psi_si.stride(0), ptr(occ_d), ptr(n_raw),
                                        stream_ptr()), "qds_density")
        self.n_raw = n_raw
        self.m_raw = float(n_raw.mean())
        self.ND_mean = float(self.N_D.mean())
        self.n_e = n_raw * (self.ND_mean / self.m_raw)
        self.e = torch.where(self.phi > 0, torch.exp(-Q_E * self.phi / self.kT), torch.ones_like(self.phi))
        self.p = self.N_A * self.e
        self.dp = torch.where(self.phi > 0, -Q_E / self.kT * self.p, torch.zeros_like(self.p))
        # the map at phi (its distance from phi says how converged phi is)
        rho = Q_E * (self.p - self.n_e)
        self.F = pp.solve(rho, self.V_p, self.V_n, x0=self.phi, rtol=1e-12)
        self.fixed_point_residual = float((self.F - self.phi).abs().max())
        scale = max(abs(self.V_p), abs(self.V_n), float(self.phi.abs().max()), 1e-3)
        if self.fixed_point_residual > 1e-6 * scale:
            warnings.warn(f"phi is {self.fixed_point_residual:.2e} V from a fixed point of the self-consistent map: "
                          f"the adjoint gives the gradient of a fixed point", RuntimeWarning, stacklevel=2)
        # Sternheimer set: states that carry occupation
        self.active = [s for s in range(self.k) if self.occ[s] > occ_floor * self.occ.max()]
        self.D = _pair_weights(self.lam, self.occ, self.docc)
        self._st = None
        self.sternheimer_iterations = []

    # -- pieces ---------------------------------------------------------------
    def _interp_T(self, gV):
        """Transpose of the nodal -> quadrature interpolation: (E, Q) -> nodal."""
        sp_ = self.sp.space
        tab = sp_.tables[sp_.nq: sp_.nq + sp_.nq * sp_.npe].reshape(sp_.nq, sp_.npe)
        out = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        out.index_add_(0, sp_.conn.long().reshape(-1), (gV @ tab).reshape(-1))
        return out

    def _sternheimer_setup(self):
        if self._st is not None:
            return self._st
        sp = self.sp
        op = sp.operator(torch.float64)
        ka = len(self.active)
        Pf = self.psi[sp.old_of_new].contiguous()                       # (n_free, k)
        MP = torch.empty_like(Pf)
        op.apply_m(Pf, MP)
        lam = self.lam
        pre = sp.preconditioner(ka, torch.float64)
        sigma0 = pre.sigma
        gap = lam[min(self.k - 1, max(self.active) + 1)] - lam[0] if self.k > 1 else 0.0
        pre.set_sigma(max(sigma0, lam[0] - max(0.25 * gap, 1e-3)))
        self._st = {"op": op, "Pf": Pf, "MP": MP, "pre": pre, "ka": ka,
                    "E": torch.from_numpy(lam[self.active].copy()).to(self.dev)}
        return self._st

    def _project_res(self, R, st):
        """R - M Psi (Psi^T R): removes the computed states from a right-hand side."""
        C = kernels.gram(st["Pf"], R)
        T = torch.empty_like(R)
        kernels.block_update(st["MP"], C, T)
        R.sub_(T)

    def _project_vec(self, Z, st):
        """Z - Psi (Psi^T M Z): M-orthogonal to the computed states."""
        C = kernels.gram(st["MP"], Z)
        T = torch.empty_like(Z)
        kernels.block_update(st["Pf"], C, T)
        Z.sub_(T)

    def _sternheimer(self, Rhs):
        """Y (n_free, ka) with (H - E_s M) y_s = (I - M Psi Psi^T) r_s, y_s
        M-orthogonal to the computed states: block CG, one shift per column,
        flexible (Polak-Ribiere) for the multigrid preconditioner."""
        st = self._sternheimer_setup()
        op, pre, E = st["op"], st["pre"], st["E"]
        R = Rhs.clone()
        self._project_res(R, st)
        Y = torch.zeros_like(R)
        r0 = torch.sqrt(kernels.coldot(R, R))
        live = r0 > 0
        if not bool(live.any()):
            return Y
        Z = torch.empty_like(R)
        pre(R, Z)
        self._project_vec(Z, st)
        Pd = Z.clone()
        rz = kernels.coldot(R, Z)
        HP, MPd, AP = torch.empty_like(R), torch.empty_like(R), torch.empty_like(R)
        its = 0
        for its in range(1, 500):
            op.apply(Pd, HP, MPd)
            kernels.block_residual(HP, MPd, E, AP)                 # (H - E_s M) p_s
            self._project_res(AP, st)
            pAp = kernels.coldot(Pd, AP)
            alpha = torch.where(live & (pAp > 0), rz / torch.where(pAp > 0, pAp, 1.0), torch.zeros_like(rz))
            Y.add_(Pd * alpha)
            R_old = R.clone()
            R.sub_(AP * alpha)
            rn = torch.sqrt(kernels.coldot(R, R))
            live = live & (rn > self.rtol_st * r0)
            if not bool(live.any()):
                break
            pre(R, Z)
            self._project_vec(Z, st)
            rz_new = kernels.coldot(R, Z)
            beta = torch.where(live, kernels.coldot(Z, R - R_old) / rz, torch.zeros_like(rz))
            rz = rz_new
            Pd.mul_(beta).add_(Z)
        self.sternheimer_iterations.append(its)
        if bool(live.any()):
            warnings.warn(f"Sternheimer block CG stopped after {its} iterations above its tolerance "
                          f"{self.rtol_st:.1e}: the gradient is less accurate", RuntimeWarning, stacklevel=3)
        self._project_vec(Y, st)
        return Y

    def eigen_adjoint(self, g_raw):
        """dJ/dV (E, Q) (per eV) for a nodal gradient g_raw with respect to n_raw."""
        k = self.k
        g_raw = g_raw.contiguous()
        Gpsi = (self.psi * g_raw[:, None]).contiguous()
        A = (C0 * C0) * kernels.gram(self.psi, Gpsi).cpu().numpy()              # A_st = c0^2 sum g psi_s psi_t
        A = 0.5 * (A + A.T)
        # occupations (diagonal): d occ_s = f'_s (dE_s - dE_0)
        cdiag = np.diag(A) * self.docc
        coef = cdiag.copy()
        coef[0] -= cdiag.sum()
        gV = self.hf @ torch.from_numpy(coef).to(self.dev)
        # rotations inside the computed states
        B = A * self.D
        beta, U = np.linalg.eigh(B)
        if np.abs(beta).max() > 0:
            rot = (self.psi @ torch.from_numpy(U).to(self.dev)).contiguous()
            gV = gV + state_density_q(self.sp.space, rot) @ torch.from_numpy(beta).to(self.dev)
        # the rest of the spectrum (Sternheimer)
        if self.active:
            act = torch.as_tensor(self.active, device=self.dev)
            w = torch.from_numpy(2.0 * self.occ[self.active] * C0 * C0).to(self.dev)
            Rhs = (Gpsi[:, act] * w)[self.sp.old_of_new].contiguous()
            Yf = self._sternheimer(Rhs)
            Y = torch.zeros((self.n, len(self.active)), dtype=torch.float64, device=self.dev)
            Y[self.sp.old_of_new] = Yf
            Ps = self.psi[:, act]
            ny = torch.linalg.vector_norm(Y, dim=0)
            npsi = torch.linalg.vector_norm(Ps, dim=0)
            a = torch.sqrt(torch.where(ny > 0, npsi / torch.where(ny > 0, ny, 1.0), torch.ones_like(ny)))
            plus = state_density_q(self.sp.space, (Y * a + Ps / a).contiguous())
            minus = state_density_q(self.sp.space, (Y * a - Ps / a).contiguous())
            gV = gV - 0.25 * (plus - minus).sum(dim=2)
        return gV

    def _pieces(self, mu):
        """Adjoint pieces for a nodal mu: (Poisson adjoint lambda, g_rho, g_n, g_V)."""
        pp = self.pp
        lam = torch.zeros_like(mu)
        lam[pp.free_d] = pp.solve_free(mu[pp.free_d].contiguous(), rtol=1e-12)
        s = 1.0 if pp.sign == "physical" else -1.0
        g_rho = s * pp.spmv(pp.M, lam)
        g_n = -Q_E * g_rho
        # normalisation n = mean(N_D) n_raw / mean(n_raw)
        c = self.ND_mean / self.m_raw
        g_raw = c * g_n - (self.ND_mean / (self.m_raw ** 2 * self.n)) * torch.dot(g_n, self.n_raw)
        g_V = self.eigen_adjoint(g_raw)
        return lam, g_rho, g_n, g_V

    def jt(self, mu):
        """(dF/dphi)^T mu."""
        _, g_rho, _, g_V = self._pieces(mu)
        return Q_E * g_rho * self.dp - self._interp_T(g_V)

    def gradients(self, g_phi=None, g_levels=None, rtol=1e-10):
        """Gradients of J = g_phi . phi* + g_levels . E* (E in eV, k levels): dict
        with "v0_q" (E, Q, per eV), "N_A", "N_D" (nodal), "V_p", "V_n", and the
        adjoint's GMRES iterations and residual."""
        g = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        if g_phi is not None:
            g = g + torch.as_tensor(np.asarray(g_phi.cpu() if torch.is_tensor(g_phi) else g_phi,
                                               dtype=np.float64), device=self.dev)
        gE_V = None
        if g_levels is not None:
            gE = torch.as_tensor(np.asarray(g_levels, dtype=np.float64), device=self.dev)
            gE_V = self.hf @ gE                                                  # dJ/dV from the levels
            g = g - self._interp_T(gE_V)
        mu, its, rel = _gmres(lambda v: v - self.jt(v), g, rtol)
        if rel > 10 * rtol:
            warnings.warn(f"adjoint GMRES stopped at relative residual {rel:.2e} (rtol {rtol:.1e})",
                          RuntimeWarning, stacklevel=2)
        lam, g_rho, g_n, g_V = self._pieces(mu)
        pp = self.pp
        g_dir = mu - pp.spmv(pp.K, lam)
        out = {"v0_q": g_V + (gE_V if gE_V is not None else 0.0),
               "N_A": Q_E * g_rho * self.e,
               "N_D": torch.full((self.n,), float(torch.dot(g_n, self.n_e)) / (self.n * self.ND_mean),
                                 dtype=torch.float64, device=self.dev),
               "V_p": float(g_dir[pp.left_d].sum()), "V_n": float(g_dir[pp.right_d].sum()),
               "gmres_iterations": its, "gmres_relres": rel, "mu": mu}
        sync()
        return out


def fixed_point_gradients(sc, phi, V_p, V_n, N_A, N_D, k, g_phi=None, g_levels=None, rtol=1e-10, eig_tol=1e-11):
    """Gradients of J = g_phi . phi* + g_levels . E* at the self-consistent
    solution phi (nodal, V) of ``sc`` (a SelfConsistentSolver run with the same
    V_p, V_n, N_A, N_D and k = num_quantum_states).  See ``SCLinearization``."""
    lin = SCLinearization(sc, phi, V_p, V_n, N_A, N_D, k, eig_tol=eig_tol)
    out = lin.gradients(g_phi, g_levels, rtol)
    out["fixed_point_residual"] = lin.fixed_point_residual
    out["eigenvalues_eV"] = lin.lam
    out["sternheimer_iterations"] = list(lin.sternheimer_iterations)
    return out


class _SCFixedPoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0_q, N_A, N_D, V_p, V_n, sc, k, opts):
        dev = v0_q.device
        sc.v0_q = v0_q.detach().contiguous()
        NA = N_A.detach().cpu().numpy()
        ND = N_D.detach().cpu().numpy()
        res = sc.solve_self_consistent(float(V_p), float(V_n), num_quantum_states=k, N_A=NA if NA.ndim else float(NA),
                                       N_D=ND if ND.ndim else float(ND))
        if not res["converged"]:
            raise RuntimeError(f"the self-consistent loop did not reach its tolerance in {res['iterations']} "
                               f"iterations (last residual {res['convergence_history'][-1]:.3e} V): the gradient "
                               f"is defined at a fixed point only")
        phi = torch.from_numpy(res["potential"]).to(dev)
        ctx.sc, ctx.k, ctx.opts = sc, k, opts
        ctx.args = (res["potential"], float(V_p), float(V_n), NA, ND)
        ctx.scalar = (NA.ndim == 0, ND.ndim == 0)
        return phi, torch.as_tensor(res["eigenvalues_eV"], dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g_phi, g_E):
        phi, V_p, V_n, NA, ND = ctx.args
        lin = SCLinearization(ctx.sc, phi, V_p, V_n, NA, ND, ctx.k, eig_tol=ctx.opts["eig_tol"])
        out = lin.gradients(g_phi, None if g_E is None else g_E.detach().cpu().numpy(), ctx.opts["rtol"])
        dev = out["N_A"].device
        g_na = out["N_A"].sum() if ctx.scalar[0] else out["N_A"]
        g_nd = out["N_D"].sum() if ctx.scalar[1] else out["N_D"]
        return (out["v0_q"], g_na, g_nd, torch.tensor(out["V_p"], dtype=torch.float64, device=dev),
                torch.tensor(out["V_n"], dtype=torch.float64, device=dev), None, None, None)


def self_consistent_fixed_point(sc, v0_q, N_A, N_D, V_p, V_n, k, rtol=1e-10, eig_tol=1e-11):
    """(phi* nodal V, lowest k levels eV) of the self-consistent loop of ``sc`` with
    base potential v0_q (E, Q, eV), doping N_A, N_D (scalar or nodal tensors, m^-3)
    and contact voltages V_p, V_n, differentiable in all five through the adjoint
    of the fixed point (``SCLinearization``).  The loop runs with the solver's own
    mixing, iteration limit and tolerance, which must reach a fixed point (set
    ``sc.set_solver_parameters(max_iterations, tolerance)`` tightly, e.g. 1e-12 V);
    it raises otherwise."""
    dev = v0_q.device

    def t(a):
        return a if torch.is_tensor(a) else torch.as_tensor(a, dtype=torch.float64, device=dev)
    return _SCFixedPoint.apply(v0_q, t(N_A), t(N_D), t(V_p), t(V_n), sc, int(k),
                               {"rtol": rtol, "eig_tol": eig_tol})
