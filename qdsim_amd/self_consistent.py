"""Self-consistent Poisson-Schroedinger loop on the GPU.

Call surface of CythonSelfConsistentSolver
(qdsim_cython/solvers/self_consistent_solver.pyx:29-340): constructor
``(mesh, epsilon_r_func, m_star_func, potential_func, temperature=300.0)``,
``solve_self_consistent(V_p, V_n, num_quantum_states=5, N_A=1e24, N_D=1e24)``
returning the dict of :305-318, ``set_solver_parameters``, ``get_potential``,
``get_carrier_densities``, ``get_quantum_states``, ``get_convergence_info``.

Physics as in the reference loop (:165-301): linear initial ramp (:210-220);
electron potential energy V_total = V - q phi (:222-242); Fermi level
E_F = E_0 + kT/2 (:257-258); n = sum_s f_FD(E_s) |psi_s|^2 rescaled to mean N_D
(:260-277); classical holes p = N_A exp(-q phi / kT) for phi > 0 else N_A
(:279-286); rho = q (p - n) (:116-117); Poisson step; residual max |d phi| (:191).
Changes that make the loop run and converge (the reference raises AttributeError
in its first iteration, :233-242): phi is interpolated to the quadrature points
with the FE basis instead of a floor-based nearest-node lookup, the nodal charge
uses each node's own density, and the new potential is mixed linearly,
phi <- phi + alpha (phi_Poisson - phi), alpha = 0.3 (the damping of
backend/src/self_consistent.cpp:1270-1272).

``anderson=m > 0`` replaces the linear mixing by Anderson mixing over the last m
iterates (no reference counterpart): with f_i = F(phi_i) - phi_i,
gamma = argmin || f_i - sum_j gamma_j (f_j+1 - f_j) || over the history and
phi <- phi_i + alpha f_i - sum_j gamma_j ((phi_j+1 - phi_j) + alpha (f_j+1 - f_j)).
Strong charge feedback (high doping, large domains) makes the linear map's
Jacobian eigenvalues reach far below -1, where linear mixing with any fixed alpha
oscillates or diverges; Anderson mixing (a secant method, GMRES-like on the
linearised map) converges there.  The history is restarted when its
least-squares system is singular.
"""
import time
import warnings

import numpy as np
import torch

from . import _lib, kernels
from .device import ptr, stream_ptr
from .poisson import PoissonProblem
from .schrodinger import SchrodingerProblem
from .units import EV_TO_J, K_B, Q_E


def mix_step(phi, phi_new, hist, alpha, depth):
    """Next potential from phi and its Poisson image phi_new (device or host
    tensors): linear mixing phi + alpha (phi_new - phi) for depth 0, else Anderson
    mixing over ``hist`` (the last depth + 1 iterates and residuals, updated in
    place).  Contacts hold V_p, V_n in phi and phi_new alike, so they never move."""
    f = phi_new - phi
    if depth <= 0:
        return phi + alpha * f
    hist.append((phi, f))
    if len(hist) > depth + 1:
        hist.pop(0)
    if len(hist) == 1:
        return phi + alpha * f
    dX = torch.stack([hist[j + 1][0] - hist[j][0] for j in range(len(hist) - 1)], 1)
    dF = torch.stack([hist[j + 1][1] - hist[j][1] for j in range(len(hist) - 1)], 1)
    # min || f - dF gamma || by the normal equations with a relative Tikhonov
    # floor; a singular or non-finite system drops the history
    A = (dF.T @ dF).cpu().numpy()
    r = (dF.T @ f).cpu().numpy()
    reg = 1e-12 * max(float(np.trace(A)), 1e-300)
    try:
        gamma = np.linalg.solve(A + reg * np.eye(A.shape[0]), r)
    except np.linalg.LinAlgError:
        gamma = None
    if gamma is None or not np.all(np.isfinite(gamma)):
        del hist[:-1]
        return phi + alpha * f
    g = torch.from_numpy(gamma).to(device=phi.device, dtype=phi.dtype)
    return phi + alpha * f - (dX + alpha * dF) @ g


def _mean(a):
    return float(a.mean().item()) if torch.is_tensor(a) else float(a)


class SelfConsistentSolver:
    def __init__(self, mesh, epsilon_r_func, m_star_func, potential_func, temperature=300.0,
                 mixing=0.3, poisson_sign="physical", anderson=0):
        self.mesh = mesh
        self.temperature = float(temperature)
        self.max_iterations = 100
        self.tolerance = 1e-6
        self.mixing = float(mixing)
        self.anderson = int(anderson)   # Anderson history depth (0: linear mixing)
        self.warm_start_dphi = 1e-2
        self.eig_tol = 1e-6           # relative residual of each eigensolve
        self.fermi_level = 0.0
        n = mesh.num_nodes
        self.potential = np.zeros(n)
        self.electron_density = np.zeros(n)
        self.hole_density = np.zeros(n)
        self.eigenvalues = np.array([], dtype=np.float64)
        self.eigenvectors = []
        self.last_solve_time = 0.0
        self.last_iterations = 0
        self.converged = False
        self.convergence_history = []
        # "physical": -div(eps grad phi) = rho; "reference": the reference's K phi = -M rho
        self.poisson = PoissonProblem(mesh, epsilon_r_func, sign=poisson_sign)
        self.schrodinger = SchrodingerProblem(mesh, m_star_func, None)
        sp_ = self.schrodinger.space
        # base potential (eV) at the quadrature points of the Schroedinger space
        self.v0_q = sp_.sample(potential_func, "quadrature", 1.0 / EV_TO_J)
        tab_phi = sp_.tables[sp_.nq: sp_.nq + sp_.nq * sp_.npe].contiguous()
        self._phi_tab = tab_phi
        x, _ = mesh.get_nodes()
        self._x = x

    def set_solver_parameters(self, max_iterations=100, tolerance=1e-6):
        self.max_iterations = int(max_iterations)
        self.tolerance = float(tolerance)

    def _phi_at_quadrature(self, phi):
        sp_ = self.schrodinger.space
        out = torch.empty((sp_.n_elem, sp_.nq), dtype=torch.float64, device=phi.device)
        _lib.check(_lib.lib.qds_interpolate_q(sp_.npe, sp_.nq, sp_.n_elem, ptr(sp_.conn), ptr(self._phi_tab),
                                              ptr(phi), ptr(out), stream_ptr()), "qds_interpolate_q")
        return out

    def _densities(self, lam_ev, psi, phi, N_A, N_D):
        kT = K_B * self.temperature
        E = lam_ev * EV_TO_J
        self.fermi_level = float(E[0] + 0.5 * kT)
        arg = (E - self.fermi_level) / kT
        occ = np.where(arg < 50, 1.0 / (1.0 + np.exp(np.minimum(arg, 50))), 0.0)
        dev = psi.device
        occ_d = torch.from_numpy(occ).to(dev)
        n = torch.empty(psi.shape[0], dtype=torch.float64, device=dev)
        _lib.check(_lib.lib.qds_density(psi.shape[0], psi.shape[1], ptr(psi), psi.stride(0), ptr(occ_d),
                                        ptr(n), stream_ptr()), "qds_density")
        s = float(n.mean().item())
        if s > 0:
            n = n * (_mean(N_D) / s)
        NA = N_A if torch.is_tensor(N_A) else torch.full_like(phi, N_A)
        p = torch.where(phi > 0, NA * torch.exp(-Q_E * phi / kT), NA)
        return n, p

    def solve_self_consistent(self, V_p, V_n, num_quantum_states=5, N_A=1e24, N_D=1e24):
        """N_A, N_D: scalars as in the reference, or per-node arrays (a doping
        profile such as an abrupt junction); the electron density is rescaled to
        the mean donor density.  Each Schroedinger solve after the first starts
        from the previous iteration's block once the potential update max |dphi|
        is below ``warm_start_dphi`` (10 mV)."""
        t0 = time.time()
        dev = self.schrodinger.space.dev
        if not np.isscalar(N_A):
            N_A = torch.as_tensor(np.asarray(N_A, dtype=np.float64), device=dev)
        if not np.isscalar(N_D):
            N_D = torch.as_tensor(np.asarray(N_D, dtype=np.float64), device=dev)
        x = torch.from_numpy(self._x).to(dev)
        Lx = self.mesh.Lx
        phi = V_p + (V_n - V_p) * x / Lx
        self.convergence_history = []
        self.converged = False
        self.unconverged_solves = 0   # eigensolves that ended unconverged
        self.breakdown_solves = 0     # eigensolves that broke down or did not converge
        self.retried_solves = 0       # ... and were redone with the safe shift
        it = 0
        residual = float("inf")
        lam = np.array([])
        psi = None
        n = torch.zeros_like(phi)
        p = torch.zeros_like(phi)
        info = None
        hist = []
        while it < self.max_iterations and residual > self.tolerance:
            it += 1
            # Schroedinger with V_total = V - q phi  (eV: V - phi)
            self.schrodinger.set_potential_ev((self.v0_q - self._phi_at_quadrature(phi)).contiguous())
            self.schrodinger.assemble()
            # warm start once the potential moves by less than warm_start_dphi (V):
            # earlier outer iterations can move it by volts, and the previous block
            # may then miss the new lowest states
            warm = info if residual < self.warm_start_dphi else None
            lam, psi, info = self.schrodinger.solve(num_quantum_states, start=warm, tol=self.eig_tol)
            if not info["converged"]:
                self.unconverged_solves += 1
                warnings.warn(f"self-consistent iteration {it}: eigensolver did not converge "
                              f"(residuals {info['residuals']})", RuntimeWarning, stacklevel=2)
            if info.get("breakdown"):
                self.breakdown_solves += 1
            if info.get("retried"):
                self.retried_solves += 1
            # psi: M_nm-normalised -> SI normalisation (psi^T M_SI psi = 1)
            psi_si = psi * 1e9
            n, p = self._densities(lam, psi_si, phi, N_A, N_D)
            rho = Q_E * (p - n)
            phi_new = self.poisson.solve(rho, V_p, V_n, x0=phi)
            upd = self._mix(phi, phi_new, hist)
            residual = float((upd - phi).abs().max().item())
            phi = upd
            self.convergence_history.append(residual)
            if residual < self.tolerance:
                self.converged = True
                break
        self.last_iterations = it
        self.last_solve_time = time.time() - t0
        self.potential = phi.cpu().numpy()
        self.electron_density = n.cpu().numpy()
        self.hole_density = p.cpu().numpy()
        self.eigenvalues = np.asarray(lam) * EV_TO_J
        self.eigenvectors = [c.copy() for c in (psi * 1e9).cpu().numpy().T] if psi is not None else []
        return self._create_solution_dict()

    def _mix(self, phi, phi_new, hist):
        return mix_step(phi, phi_new, hist, self.mixing, self.anderson)

    def _create_solution_dict(self):
        return {
            'potential': self.potential.copy(),
            'electron_density': self.electron_density.copy(),
            'hole_density': self.hole_density.copy(),
            'eigenvalues': self.eigenvalues.copy(),
            'eigenvalues_eV': self.eigenvalues / EV_TO_J,
            'eigenvectors': [psi.copy() for psi in self.eigenvectors],
            'fermi_level': self.fermi_level,
            'fermi_level_eV': self.fermi_level / EV_TO_J,
            'converged': self.converged,
            'iterations': self.last_iterations,
            'solve_time': self.last_solve_time,
            'convergence_history': list(self.convergence_history),
        }

    def get_potential(self):
        return self.potential.copy()

    def get_carrier_densities(self):
        return self.electron_density.copy(), self.hole_density.copy()

    def get_quantum_states(self):
        return self.eigenvalues.copy(), [psi.copy() for psi in self.eigenvectors]

    def get_convergence_info(self):
        return {
            'converged': self.converged,
            'iterations': self.last_iterations,
            'final_residual': self.convergence_history[-1] if self.convergence_history else float('inf'),
            'solve_time': self.last_solve_time,
            'convergence_history': list(self.convergence_history),
        }
