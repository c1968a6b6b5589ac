"""Hellmann-Feynman adjoint vs central finite differences and an FP64 oracle gradient."""
import numpy as np
import pytest
import torch

from oracle import eig_ref, fem_ref

pytestmark = pytest.mark.gpu

M_E = 9.1093837015e-31
L = 40.0
CELLS = 12


def _gates(theta, x, y, lib):
    # theta = [a_g, cx_g, cy_g] * n_gates, sigma fixed per gate
    v = 0.0
    for g in range(len(theta) // 3):
        a, cx, cy = theta[3 * g], theta[3 * g + 1], theta[3 * g + 2]
        v = v + a * lib.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 6.0 ** 2))
    return v


def _problem():
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.schrodinger import SchrodingerProblem
    mesh = SimpleMesh(CELLS + 1, CELLS + 1, L * 1e-9, L * 1e-9, 2)
    prob = SchrodingerProblem(mesh, lambda x, y: 0.067 * M_E + 0 * x, None)
    return prob


def test_gradients_fd_and_oracle():
    from qdsim_amd.adjoint import eigenvalues
    prob = _problem()
    xq, yq = prob.space.quad_points()
    theta0 = np.array([-0.25, 18.0, 21.0, -0.1, 25.0, 15.0])
    target = np.array([-0.15, -0.05, 0.0])
    k = 3

    def loss_t(th):
        v_q = _gates(th, xq, yq, torch)
        E = eigenvalues(prob, v_q, k, mode="fp64", tol=1e-10)
        return ((E - torch.as_tensor(target, device=E.device)) ** 2).sum(), E

    th = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
    loss, E = loss_t(th)
    loss.backward()
    g = th.grad.cpu().numpy()
    # central differences, step 1e-6
    h = 1e-6
    fd = np.zeros_like(theta0)
    for i in range(len(theta0)):
        tp = torch.tensor(theta0, dtype=torch.float64, device="cuda"); tp[i] += h
        tm = torch.tensor(theta0, dtype=torch.float64, device="cuda"); tm[i] -= h
        with torch.no_grad():
            fd[i] = (loss_t(tp)[0].item() - loss_t(tm)[0].item()) / (2 * h)
    print("adjoint", g, "fd", fd)
    assert np.allclose(g, fd, rtol=1e-4, atol=1e-7)
    # FP64 oracle gradient: eigenvectors from scipy, Hellmann-Feynman by oracle quadrature
    vfun = lambda x, y: _gates(theta0, x, y, np)
    r = eig_ref.solve_closed(CELLS + 1, CELLS + 1, L, L, 2, lambda x, y: 0.067 + 0 * x, vfun, k)
    assert np.allclose(E.detach().cpu().numpy(), r["vals"], rtol=1e-9)
    x, y, el = r["x"], r["y"], r["el"]
    vx, vy = x[el[:, :3]], y[el[:, :3]]
    P, W = fem_ref.rule(6)
    area = 0.5 * np.abs((vx[:, 1] - vx[:, 0]) * (vy[:, 2] - vy[:, 0]) - (vx[:, 2] - vx[:, 0]) * (vy[:, 1] - vy[:, 0]))
    phi = np.stack(fem_ref.shape_values(2, P[:, 0], P[:, 1], P[:, 2]), -1)      # Q x 6
    psi_q = np.einsum("qa,eak->eqk", phi, r["vecs"][el])                        # E x Q x k
    rho = (W[None, :, None] * area[:, None, None]) * psi_q ** 2
    xqo, yqo = fem_ref.quad_points(vx, vy, 6)
    dEdv = 2 * (r["vals"] - target)                                               # dloss/dE
    go = np.zeros_like(theta0)
    eps = 1e-7
    for i in range(len(theta0)):
        tp = theta0.copy(); tp[i] += eps
        tm = theta0.copy(); tm[i] -= eps
        dv = (_gates(tp, xqo, yqo, np) - _gates(tm, xqo, yqo, np)) / (2 * eps)   # dV/dtheta at quad pts
        go[i] = np.sum(np.einsum("eqk,eq->k", rho, dv) * dEdv)
    print("oracle", go)
    assert np.allclose(g, go, rtol=1e-6, atol=1e-9)


def test_nodal_gradient_sums_to_state_norm():
    """sum_i dE_k/dV_i = 1 for a constant shift (psi^T M psi = 1)."""
    from qdsim_amd.adjoint import eigenvalues, nodal_to_quadrature
    prob = _problem()
    n = prob.space.n
    v = torch.zeros(n, dtype=torch.float64, device="cuda", requires_grad=True)
    xq, yq = prob.space.quad_points()
    base = _gates(np.array([-0.2, 20.0, 20.0]), xq, yq, torch)
    E = eigenvalues(prob, base + nodal_to_quadrature(prob.space, v), 2, mode="fp64", tol=1e-10)
    E[0].backward()
    assert abs(v.grad.sum().item() - 1.0) < 1e-9
