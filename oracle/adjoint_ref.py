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
