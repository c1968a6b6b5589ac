"""FP64 reference gradients of the gate-design loss (oracle; test use only).

No reference counterpart: the reference computes no gradients.  The loss is
sum_k (E_k - E*_k)^2 over the lowest k levels of the closed pencil with the
potential V = sum_g a_g exp(-|r - c_g|^2 / 2 s_g^2) (theta = (a_g, cx_g, cy_g) per
gate, widths s_g fixed).  Eigenpairs from eig_ref (scipy shift-invert), then
Hellmann-Feynman by the oracle's own quadrature:
    dE_k / dtheta_j = sum_e sum_q w_q |J_e| / 2  psi_k(x_q)^2  dV/dtheta_j(x_q)
with psi_k M-normalised and dV/dtheta_j in closed form.  Pinned by central
differences of the eig_ref eigenvalues (tests/test_oracle_pinning.py).
"""
import numpy as np

from . import eig_ref, fem_ref


def gate_v(theta, sigma, x, y):
    """V (eV) at points (nm)."""
    v = np.zeros_like(np.asarray(x, dtype=np.float64))
    for g in range(len(sigma)):
        a, cx, cy = theta[3 * g], theta[3 * g + 1], theta[3 * g + 2]
        v = v + a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sigma[g] ** 2))
    return v


def gate_dv(theta, sigma, x, y):
    """dV/dtheta (len(theta), *x.shape): d/da = e, d/dcx = a e (x - cx) / s^2,
    d/dcy = a e (y - cy) / s^2 with e the Gaussian."""
    out = np.zeros((len(theta),) + np.shape(x))
    for g in range(len(sigma)):
        a, cx, cy = theta[3 * g], theta[3 * g + 1], theta[3 * g + 2]
        s2 = sigma[g] ** 2
        e = np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * s2))
        out[3 * g] = e
        out[3 * g + 1] = a * e * (x - cx) / s2
        out[3 * g + 2] = a * e * (y - cy) / s2
    return out


def loss_and_gradient(cells, order, L_nm, m_rel, theta, sigma, targets):
    """(loss, gradient (len(theta),), eigenvalues (k,)) on a cells x cells order-p
    mesh of an L x L nm square with constant m* = m_rel m0."""
    k = len(targets)
    r = eig_ref.solve_closed(cells + 1, cells + 1, L_nm, L_nm, order, lambda x, y: m_rel + 0 * x,
                             lambda x, y: gate_v(theta, sigma, x, y), k)
    x, y, el = r["x"], r["y"], r["el"]
    vx, vy = x[el[:, :3]], y[el[:, :3]]
    deg = 2 * order + 2
    P, W = fem_ref.rule(deg)
    area = 0.5 * np.abs((vx[:, 1] - vx[:, 0]) * (vy[:, 2] - vy[:, 0]) - (vx[:, 2] - vx[:, 0]) * (vy[:, 1] - vy[:, 0]))
    phi = np.stack(fem_ref.shape_values(order, P[:, 0], P[:, 1], P[:, 2]), -1)     # Q x npe
    psi_q = np.einsum("qa,eak->eqk", phi, r["vecs"][el])                           # E x Q x k
    rho = (W[None, :, None] * area[:, None, None]) * psi_q ** 2                    # w |J|/2 psi^2
    xq, yq = fem_ref.quad_points(vx, vy, deg)
    dv = gate_dv(theta, sigma, xq, yq)                                             # p x E x Q
    dE = np.einsum("eqk,peq->kp", rho, dv)                                         # k x p
    res = r["vals"] - np.asarray(targets)
    return float(np.sum(res ** 2)), 2.0 * res @ dE, r["vals"]


# ---- gate (contact) voltages and doping through the Poisson step ----------------
Q_E = 1.602176634e-19


def doping_rho(theta, x_nm, width=5.0):
    """rho (C/m^3) of a smooth junction: donors N_D0 s(x) and acceptors N_A0 (1 - s(x)),
    s = 1 / (1 + exp(-(x - x_j) / width)); theta = (V_p, V_n, N_D0, N_A0, x_j) with
    densities in 1e24 m^-3 and x_j in nm.  Returns (rho, d rho / d theta[2:])."""
    nd, na, xj = theta[2], theta[3], theta[4]
    s = 1.0 / (1.0 + np.exp(-(x_nm - xj) / width))
    c = Q_E * 1e24
    rho = c * (nd * s - na * (1.0 - s))
    ds = -s * (1.0 - s) / width
    return rho, np.stack([c * s, -c * (1.0 - s), c * (nd + na) * ds])


def electrostatic_loss_and_gradient(nx, ny, Lx_nm, Ly_nm, eps_r, m_rel, v0_func, theta, targets,
                                    sign="physical"):
    """J = sum_k (E_k - E*_k)^2 for the lowest levels of V0 - q phi on a P1 mesh of
    nx x ny nodes, phi from poisson_ref with contacts V_p, V_n and the charge of
    ``doping_rho``; gradient with respect to theta = (V_p, V_n, N_D0, N_A0, x_j) by
    Hellmann-Feynman (oracle quadrature), the transpose of the P1 interpolation and
    one adjoint solve with the Poisson matrix (spsolve).  Returns (J, grad, E)."""
    import scipy.sparse.linalg as spla
    from . import poisson_ref
    k = len(targets)
    x_nm, y_nm, el = poisson_ref.poisson_matrices(nx, ny, Lx_nm, Ly_nm, lambda x, y: eps_r + 0 * x)[:3]
    _, _, _, K, M = poisson_ref.poisson_matrices(nx, ny, Lx_nm * 1e-9, Ly_nm * 1e-9, lambda x, y: eps_r + 0 * x)
    rho, drho = doping_rho(theta, x_nm)
    phi, _, _ = poisson_ref.solve(nx, ny, Lx_nm * 1e-9, Ly_nm * 1e-9, lambda x, y: eps_r + 0 * x, rho, theta[0],
                                  theta[1], sign=sign)
    deg = 4                                           # the P1 rule of eig_ref (2p + 2)
    P, W = fem_ref.rule(deg)
    N = np.stack(fem_ref.shape_values(1, P[:, 0], P[:, 1], P[:, 2]), -1)           # Q x 3
    phi_q = phi[el] @ N.T                                                          # E x Q

    def v_total(x, y):
        if np.shape(x) == phi_q.shape:
            return v0_func(x, y) - phi_q
        return v0_func(x, y) - phi                    # nodal (the oracle's shift estimate)
    r = eig_ref.solve_closed(nx, ny, Lx_nm, Ly_nm, 1, lambda x, y: m_rel + 0 * x, v_total, k)
    vx, vy = x_nm[el], y_nm[el]
    area = 0.5 * np.abs((vx[:, 1] - vx[:, 0]) * (vy[:, 2] - vy[:, 0]) - (vx[:, 2] - vx[:, 0]) * (vy[:, 1] - vy[:, 0]))
    psi_q = np.einsum("qa,eak->eqk", N, r["vecs"][el])
    hf = (W[None, :, None] * area[:, None, None]) * psi_q ** 2                     # dE_k / dV(x_q)
    res = r["vals"] - np.asarray(targets)
    gV = hf @ (2.0 * res)                                                          # dJ / dV(x_q), E x Q
    g_phi = np.zeros(len(x_nm))
    np.add.at(g_phi, el, -(gV @ N))                                                # V - q phi, transposed
    tol = 1e-9 * Lx_nm
    left = np.abs(x_nm) < tol
    right = np.abs(x_nm - Lx_nm) < tol
    free = ~(left | right)
    lam = np.zeros(len(x_nm))
    lam[free] = spla.spsolve(K[free][:, free].tocsc(), g_phi[free])
    s = 1.0 if sign == "physical" else -1.0
    g_rho = s * (M @ lam)
    g_dir = g_phi - K @ lam
    grad = np.concatenate([[g_dir[left].sum(), g_dir[right].sum()], drho @ g_rho])
    return float(np.sum(res ** 2)), grad, r["vals"]
