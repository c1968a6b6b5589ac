"""One outer iteration of the self-consistent Poisson-Schroedinger loop (oracle;
test use only).

Restates qdsim_cython/solvers/self_consistent_solver.pyx:165-301 from its first
iteration on a P1 mesh, composed of the pinned oracles ``eig_ref`` (closed
eigenproblem) and ``poisson_ref`` (Poisson with the chosen sign):

* initial potential: the linear ramp phi0 = V_p + (V_n - V_p) x / Lx (:210-220);
* electron potential energy V - q phi0 (:222-242); the FE interpolant of the
  ramp is the ramp itself, so no node lookup is involved;
* Fermi level E_0 + kT/2 (:257-258), occupations 1 / (1 + exp((E - E_F)/kT)),
  zero when the exponent reaches 50 (:260-270);
* n = sum_s occ_s psi_s^2 at the nodes, rescaled to mean N_D (:272-277);
* classical holes p = N_A exp(-q phi0 / kT) where phi0 > 0, else N_A (:279-286);
* rho = q (p - n) (:116-117), one Poisson solve, linear mixing
  phi1 = phi0 + alpha (phi_P - phi0) (backend/src/self_consistent.cpp:1270-1272).

N_A, N_D: scalars or per-node arrays (the mean of N_D is the target of the
rescaling).  Pinned by composition: eig_ref and poisson_ref carry their own
pins (tests/test_oracle_pinning.py), and the steps in between are the formulas
above; alpha = 0 returns phi0 exactly.
"""
import numpy as np

from . import eig_ref, poisson_ref

K_B = 1.380649e-23
Q_E = 1.602176634e-19


def one_iteration(nx, ny, Lx_nm, Ly_nm, eps_r, mstar_func, v_func, V_p, V_n, k, N_A, N_D,
                  temperature=300.0, mixing=0.3, sign="physical"):
    """mstar_func(x_nm, y_nm) -> m*/m0, v_func(x_nm, y_nm) -> eV.  Returns a dict with
    the eigenvalues (eV), n, p, rho, the Poisson potential and the mixed potential
    (nodal, full numbering, SI)."""

    def ramp_nm(x):
        return V_p + (V_n - V_p) * x / Lx_nm

    def v_total(x, y):
        return v_func(x, y) - ramp_nm(x)          # eV: q phi in eV is phi in volts

    m = eig_ref.solve_closed(nx, ny, Lx_nm, Ly_nm, 1, mstar_func, v_total, k)
    x_nm = m["x"]
    phi0 = ramp_nm(x_nm)
    kT = K_B * temperature
    E = m["vals"] * eig_ref.EV
    fermi = E[0] + 0.5 * kT
    arg = (E - fermi) / kT
    occ = np.where(arg < 50, 1.0 / (1.0 + np.exp(np.minimum(arg, 50))), 0.0)
    n = (m["vecs"] ** 2) @ occ
    n = n * (np.mean(N_D) / n.mean())
    NA = np.broadcast_to(np.asarray(N_A, dtype=np.float64), phi0.shape)
    p = np.where(phi0 > 0, NA * np.exp(-Q_E * phi0 / kT), NA)
    rho = Q_E * (p - n)
    phi_p, _, _ = poisson_ref.solve(nx, ny, Lx_nm * 1e-9, Ly_nm * 1e-9, lambda x, y: eps_r + 0 * x, rho, V_p, V_n,
                                    sign=sign)
    phi1 = phi0 + mixing * (phi_p - phi0)
    return {"vals": m["vals"], "n": n, "p": p, "rho": rho, "phi0": phi0, "phi_poisson": phi_p, "phi": phi1,
            "residual": float(np.abs(phi1 - phi0).max()), "fermi_level": fermi}
