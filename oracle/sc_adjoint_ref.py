"""Self-consistent fixed point and its gradients, dense FP64 (oracle; test use only).

No reference counterpart for the gradients (the reference computes none).  The
map is the outer iteration of qdsim_cython/solvers/self_consistent_solver.pyx:
165-301 as ``sc_ref.one_iteration`` states it, for a general potential phi
instead of the initial ramp:

    V = v0 - phi (eV, at the quadrature points: phi interpolated with P1),
    (E_s, psi_s) lowest k of H(V) psi = E M psi (dense generalized eigh, all states),
    E_F = E_0 + kT/2, occ_s = 1 / (1 + exp((E_s - E_F)/kT)) (0 from exponent 50),
    n = mean(N_D) n_raw / mean(n_raw),  n_raw = sum_s occ_s psi_s^2 (SI scale),
    p = N_A exp(-q phi / kT) where phi > 0, else N_A,   rho = q (p - n),
    F(phi) = Poisson(rho; V_p at x = 0, V_n at x = Lx)   (poisson_ref, chosen sign).

The fixed point phi* = F(phi*) is found by Newton's method with the dense
Jacobian, to machine precision.  The Jacobian uses first-order perturbation
theory over the FULL spectrum of the dense pencil:
    dE_s = <psi_s, dH psi_s>,
    dn_raw = sum_s [docc_s psi_s^2 + 2 occ_s psi_s dpsi_s],
    dpsi_s = sum_{t != s} psi_t <psi_t, dH psi_s> / (E_s - E_t),
with the pairs of computed states combined into divided differences
(occ_s - occ_t) / (E_s - E_t) (finite through degenerate pairs), and
<psi_s, dH psi_t> for a change dV at one quadrature point equal to
w_q |J_e| psi_s(x_q) psi_t(x_q) (the Hellmann-Feynman weights of adjoint_ref).
Gradients of J = g_phi . phi* + g_E . E* follow by the adjoint of the fixed
point, mu = (I - dF/dphi)^-T (g_phi + (dE/dphi)^T g_E).

Pinned by central differences of the Newton-converged fixed point with respect to
every parameter kind (tests/test_oracle_pinning.py), and the Jacobian by central
differences of F.  The map itself is ``sc_ref.one_iteration``'s (checked against
it at the ramp).
"""
import numpy as np
import scipy.linalg as sla

from . import eig_ref, fem_ref, mesh_ref, poisson_ref

K_B = 1.380649e-23
Q_E = 1.602176634e-19
EV = eig_ref.EV


class SCProblem:
    """P1 mesh of nx x ny nodes on Lx x Ly nm; constant eps_r and m* (m0 units);
    v0_q: base potential (eV) at the quadrature points (E, Q)."""

    def __init__(self, nx, ny, Lx_nm, Ly_nm, eps_r, m_rel, k, temperature=300.0, sign="physical"):
        self.nx, self.ny, self.Lx, self.Ly, self.k = nx, ny, Lx_nm, Ly_nm, k
        self.sign = sign
        self.kT = K_B * temperature
        x, y, el = mesh_ref.reference_mesh(nx, ny, Lx_nm, Ly_nm, 1)
        self.x, self.y, self.el = x, y, el
        self.n = len(x)
        vx, vy = x[el], y[el]
        self.deg = 4
        P, W = fem_ref.rule(self.deg)
        self.N = np.stack(fem_ref.shape_values(1, P[:, 0], P[:, 1], P[:, 2]), -1)      # Q x 3
        self.xq, self.yq = fem_ref.quad_points(vx, vy, self.deg)                       # E x Q
        area = 0.5 * np.abs((vx[:, 1] - vx[:, 0]) * (vy[:, 2] - vy[:, 0]) - (vx[:, 2] - vx[:, 0]) * (vy[:, 1] - vy[:, 0]))
        self.wq = W[None, :] * area[:, None]                                            # E x Q
        c_q = eig_ref.KIN / (m_rel * eig_ref.M_E) * np.ones_like(self.xq)
        self._vx, self._vy, self._c_q = vx, vy, c_q
        Kk, _, Mm = fem_ref.element_matrices(1, vx, vy, self.deg, c_q=c_q, v_q=np.zeros_like(self.xq))
        self.Kkin = fem_ref.assemble(el, self.n, Kk)
        self.Ms = fem_ref.assemble(el, self.n, Mm)
        self.sfree = eig_ref.free_dofs(x, y, el, Lx_nm, Ly_nm)
        # Poisson (SI), contacts at x = 0 and x = Lx
        _, _, _, self.KP, self.MP = poisson_ref.poisson_matrices(nx, ny, Lx_nm * 1e-9, Ly_nm * 1e-9,
                                                                 lambda a, b: eps_r + 0 * a)
        tol = 1e-9 * Lx_nm
        self.left = np.abs(x) < tol
        self.right = np.abs(x - Lx_nm) < tol
        self.pfree = np.nonzero(~(self.left | self.right))[0]
        KP = self.KP.toarray()
        self._Kff_inv = np.linalg.inv(KP[np.ix_(self.pfree, self.pfree)])
        self._KP = KP
        self._s = 1.0 if sign == "physical" else -1.0
        # interpolation nodal -> quadrature, (E*Q) x n
        E, Q = self.xq.shape
        self.interp = np.zeros((E * Q, self.n))
        for a in range(3):
            np.add.at(self.interp, (np.arange(E * Q), np.repeat(el[:, a], Q)), np.tile(self.N[:, a], E))

    # -- pieces -------------------------------------------------------------
    def phi_q(self, phi):
        return (self.interp @ phi).reshape(self.xq.shape)

    def states(self, v_q):
        """All eigenpairs of the free-DOF pencil for V = v_q (eV); vectors in full
        numbering, M-normalised (nm units)."""
        _, Vm, _ = fem_ref.element_matrices(1, self._vx, self._vy, self.deg, c_q=None, v_q=v_q)
        H = (self.Kkin + fem_ref.assemble(self.el, self.n, Vm)).toarray()
        M = self.Ms.toarray()
        f = self.sfree
        vals, vecs = sla.eigh(H[np.ix_(f, f)], M[np.ix_(f, f)])
        full = np.zeros((self.n, len(vals)))
        full[f] = vecs
        return vals, full

    def poisson(self, rho, V_p, V_n):
        g = np.where(self.left, V_p, np.where(self.right, V_n, 0.0))
        f = self.pfree
        rhs = self._s * (self.MP @ rho) - self._KP @ g
        phi = g.copy()
        phi[f] = self._Kff_inv @ rhs[f]
        return phi

    def occupations(self, E):
        EJ = E * EV
        fermi = EJ[0] + 0.5 * self.kT
        arg = (EJ - fermi) / self.kT
        occ = np.where(arg < 50, 1.0 / (1.0 + np.exp(np.minimum(arg, 50))), 0.0)
        docc = np.where(arg < 50, -occ * (1.0 - occ), 0.0) * EV / self.kT     # d occ / d(E_s - E_F), per eV
        return occ, docc

    def evaluate(self, phi, v0_q, V_p, V_n, N_A, N_D):
        """F(phi) and everything the derivatives need."""
        NA = np.broadcast_to(np.asarray(N_A, dtype=np.float64), (self.n,))
        ND = np.broadcast_to(np.asarray(N_D, dtype=np.float64), (self.n,))
        v_q = v0_q - self.phi_q(phi)
        vals, vecs = self.states(v_q)
        k = self.k
        E = vals[:k]
        occ, docc = self.occupations(E)
        psi_si = vecs[:, :k] * 1e9
        n_raw = (psi_si ** 2) @ occ
        n = n_raw * (ND.mean() / n_raw.mean())
        e = np.where(phi > 0, np.exp(-Q_E * phi / self.kT), 1.0)
        p = NA * e
        rho = Q_E * (p - n)
        return {"F": self.poisson(rho, V_p, V_n), "vals": vals, "vecs": vecs, "E": E, "occ": occ, "docc": docc,
                "n_raw": n_raw, "n": n, "p": p, "e": e, "rho": rho, "NA": NA, "ND": ND, "phi": phi}

    # -- derivatives ----------------------------------------------------------
    def dnraw_dv(self, st):
        """d n_raw / d v_q: (n, E*Q)."""
        k = self.k
        vals, vecs = st["vals"], st["vecs"]
        occ, docc = st["occ"], st["docc"]
        nall = len(vals)
        E = vals
        D = np.zeros((k, nall))
        for s in range(k):
            for t in range(nall):
                if t == s:
                    continue
                if t < k:
                    dE = E[s] - E[t]
                    if abs(dE) > 1e-9 * max(1.0, abs(E[s])):
                        D[s, t] = (occ[s] - occ[t]) / dE
                    else:
                        D[s, t] = 0.5 * (docc[s] + docc[t])
                else:
                    D[s, t] = 2.0 * occ[s] / (E[s] - E[t])
        for s in range(k):
            D[s, s] = docc[s]
        psi = vecs * 1e9                                        # SI scale for n_raw
        psiq = (self.interp @ vecs).reshape(-1, nall)           # E*Q x nall (nm normalisation, as H)
        w = self.wq.reshape(-1)
        X1 = psi[:, :k, None] * psi[:, None, :] * D[None]       # n x k x nall
        X2 = w[:, None, None] * psiq[:, :k, None] * psiq[:, None, :]
        A = np.einsum("xst,est->xe", X1, X2)
        # the Fermi level moves with E_0: docc_s (dE_s - dE_0)
        a0 = (psi[:, :k] ** 2) @ docc                            # n
        A -= a0[:, None] * (w * psiq[:, 0] ** 2)[None, :]
        return A

    def hf(self, st):
        """d E_s / d v_q: (k, E*Q)."""
        psiq = (self.interp @ st["vecs"][:, :self.k]).reshape(-1, self.k)
        return (self.wq.reshape(-1)[:, None] * psiq ** 2).T

    def jacobians(self, st):
        """Dense dF/dphi (n x n), dF/dv0 (n x EQ), dF/dN_A (n x n), dF/dN_D (n x n),
        dF/dV_p, dF/dV_n (n,), dE/dphi (k x n), dE/dv0 (k x EQ)."""
        n = self.n
        f = self.pfree
        G = np.zeros((n, n))                                     # d phi_P / d rho
        G[f] = self._s * (self._Kff_inv @ self.MP.toarray()[f])
        Araw = self.dnraw_dv(st)
        nr = st["n_raw"]
        c = st["ND"].mean() / nr.mean()
        An = c * (Araw - np.outer(nr, Araw.mean(0)) / nr.mean())   # d n / d v_q
        Dp = np.where(st["phi"] > 0, -Q_E / self.kT * st["p"], 0.0)
        J = Q_E * G @ (np.diag(Dp) + An @ self.interp)
        Fv = -Q_E * G @ An
        FNA = Q_E * G * st["e"][None, :]
        FND = -Q_E * np.outer(G @ (st["n"] / st["ND"].mean()), np.ones(n)) / n
        gl = np.where(self.left, 1.0, 0.0)
        gr = np.where(self.right, 1.0, 0.0)
        FVp, FVn = gl.copy(), gr.copy()
        FVp[f] = -self._Kff_inv @ (self._KP @ gl)[f]
        FVn[f] = -self._Kff_inv @ (self._KP @ gr)[f]
        HF = self.hf(st)
        return {"J": J, "Fv": Fv, "FNA": FNA, "FND": FND, "FVp": FVp, "FVn": FVn,
                "Ephi": -HF @ self.interp, "Ev": HF}

    # -- fixed point ---------------------------------------------------------
    def fixed_point(self, v0_q, V_p, V_n, N_A, N_D, phi0=None, tol=1e-15, maxiter=60):
        """Newton on the non-contact nodes; returns (phi*, state at phi*, iterations)."""
        phi = (V_p + (V_n - V_p) * self.x / self.Lx) if phi0 is None else np.array(phi0, dtype=np.float64)
        f = self.pfree
        scale = max(abs(V_p), abs(V_n), 1e-3)
        for it in range(1, maxiter + 1):
            st = self.evaluate(phi, v0_q, V_p, V_n, N_A, N_D)
            r = phi - st["F"]
            if np.abs(r).max() <= tol * scale * 10 and it > 1:
                return phi, st, it
            J = self.jacobians(st)["J"]
            A = np.eye(len(f)) - J[np.ix_(f, f)]
            phi = phi.copy()
            phi[f] -= np.linalg.solve(A, r[f])
            phi[self.left], phi[self.right] = V_p, V_n
        st = self.evaluate(phi, v0_q, V_p, V_n, N_A, N_D)
        if np.abs(phi - st["F"]).max() > 1e-10 * scale:
            raise RuntimeError("Newton did not converge")
        return phi, st, maxiter

    def gradients(self, st, g_phi, g_E):
        """Reverse-mode gradients of J = g_phi . phi* + g_E . E* at the fixed point
        described by ``st``: dict with v0_q (E, Q), N_A, N_D (n), V_p, V_n."""
        jac = self.jacobians(st)
        f = self.pfree
        g = np.asarray(g_phi, dtype=np.float64) + jac["Ephi"].T @ np.asarray(g_E, dtype=np.float64)
        # phi* = F(phi*) holds on every node (contacts: F = V); only the non-contact
        # rows of F depend on phi, so mu = (I - J^T)^-1 g with J's contact rows zero
        mu = np.linalg.solve(np.eye(self.n) - jac["J"].T, g)
        del f
        return {"v0_q": (jac["Fv"].T @ mu + jac["Ev"].T @ g_E).reshape(self.xq.shape),
                "N_A": jac["FNA"].T @ mu, "N_D": jac["FND"].T @ mu,
                "V_p": float(jac["FVp"] @ mu), "V_n": float(jac["FVn"] @ mu), "mu": mu}
