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


@pytest.mark.parametrize("sign", ["physical", "reference"])
def test_gate_and_doping_gradient_through_poisson(sign):
    """Levels of V0 - q phi with phi from the Poisson step (contact voltages V_p, V_n
    and a smooth junction's doping as parameters): the autograd gradient (Hellmann-
    Feynman, FE interpolation transposed, adjoint CG) against the FP64 oracle gradient
    and central differences with step 1e-6."""
    from oracle import adjoint_ref
    from qdsim_amd.adjoint import electrostatic_levels
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.poisson import PoissonProblem
    from qdsim_amd.schrodinger import SchrodingerProblem
    nx, ny, Lx, Ly = 33, 17, 80.0, 40.0
    EVJ = 1.602176634e-19

    def v0(x, y):
        return -0.2 * np.exp(-((x - 40.0) ** 2 + (y - 20.0) ** 2) / (2 * 8.0 ** 2))
    theta0 = np.array([0.05, -0.1, 0.02, 0.015, 43.0])
    targets = np.array([-0.2, -0.1, -0.05])
    print(f"gate/doping gradient, sign={sign}, P1 {nx - 1}x{ny - 1} cells", flush=True)
    mesh = SimpleMesh(nx, ny, Lx * 1e-9, Ly * 1e-9, 1)
    prob = SchrodingerProblem(mesh, lambda x, y: 0.067 * M_E + 0 * x, None)
    pp = PoissonProblem(mesh, lambda x, y: 13.9 + 0 * x, sign=sign)
    v0_q = prob.space.sample(lambda x, y: v0(x * 1e9, y * 1e9) * EVJ, "quadrature", 1.0 / EVJ)
    x_nm = torch.from_numpy(mesh.get_nodes()[0] * 1e9).to("cuda")
    tgt = torch.as_tensor(targets, device="cuda")

    def loss_t(th):
        s = 1.0 / (1.0 + torch.exp(-(x_nm - th[4]) / 5.0))
        rho = 1.602176634e-19 * 1e24 * (th[2] * s - th[3] * (1.0 - s))
        E = electrostatic_levels(prob, pp, v0_q, rho, th[0], th[1], 3, mode="fp64", tol=1e-10)
        return ((E - tgt) ** 2).sum(), E

    th = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
    loss, E = loss_t(th)
    loss.backward()
    g = th.grad.cpu().numpy()
    J, go, Eo = adjoint_ref.electrostatic_loss_and_gradient(nx, ny, Lx, Ly, 13.9, 0.067, v0, theta0, targets, sign)
    print("E", E.detach().cpu().numpy(), "oracle", Eo, flush=True)
    print("grad", g, "oracle", go, flush=True)
    assert np.allclose(E.detach().cpu().numpy(), Eo, rtol=1e-9, atol=1e-12)
    assert abs(loss.item() - J) <= 1e-9 * J
    assert np.allclose(g, go, rtol=1e-6, atol=1e-9 * np.abs(go).max())
    for j in (0, 2, 4):
        h = 1e-6 * max(1.0, abs(theta0[j]))
        tp = torch.tensor(theta0, dtype=torch.float64, device="cuda"); tp[j] += h
        tm = torch.tensor(theta0, dtype=torch.float64, device="cuda"); tm[j] -= h
        with torch.no_grad():
            fd = (loss_t(tp)[0].item() - loss_t(tm)[0].item()) / (2 * h)
        print("param", j, "adjoint", g[j], "fd", fd, flush=True)
        assert abs(g[j] - fd) <= 1e-4 * abs(fd) + 1e-9
