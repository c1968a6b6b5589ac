"""Gradients through the self-consistent fixed point (qdsim_amd.sc_adjoint) against
the dense FP64 oracle (oracle/sc_adjoint_ref.py: Newton-converged fixed point,
full-spectrum perturbation theory, pinned by central differences).

The loop of self_consistent_solver.pyx:165-301 with linear mixing is run to a
fixed point (1e-13 V) on the GPU, then one adjoint: GMRES on (I - dF/dphi)^T with
each product an adjoint Poisson solve and the adjoint of the eigenstates' density
response (Hellmann-Feynman occupations, pair rotations, Sternheimer solves for the
states not computed).  Tolerances: eigensolves 1e-11, Sternheimer 1e-11, GMRES
1e-10, so the gradients agree with the oracle's to ~1e-8; the bar is 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.602176634e-19
M_E = 9.1093837015e-31


def _v_nm(x, y):
    return -0.15 * np.exp(-((x - 40.0) ** 2 + (y - 25.0) ** 2) / (2 * 10.0 ** 2))


@pytest.mark.parametrize("doping", ["uniform", "junction"])
def test_sc_fixed_point_gradients_match_oracle(doping):
    from oracle import sc_adjoint_ref
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.sc_adjoint import fixed_point_gradients
    from qdsim_amd.self_consistent import SelfConsistentSolver
    nx, ny, Lx, Ly, k = 17, 11, 80.0, 50.0, 4
    V_p, V_n = -0.05, 0.2
    P = sc_adjoint_ref.SCProblem(nx, ny, Lx, Ly, 12.9, 0.067, k)
    if doping == "uniform":
        NA, ND = np.full(P.n, 2e22), np.full(P.n, 3e22)
    else:
        left = P.x < 37.0
        NA, ND = np.where(left, 4e22, 1e21), np.where(left, 1e21, 5e22)
    print(f"SC fixed-point gradients, P1 {nx - 1}x{ny - 1} cells, k={k}, {doping} doping", flush=True)
    phi_ref, st, _ = P.fixed_point(_v_nm(P.xq, P.yq), V_p, V_n, NA, ND)
    rng = np.random.default_rng(7)
    g_phi, g_E = rng.standard_normal(P.n), rng.standard_normal(k)
    ref = P.gradients(st, g_phi, g_E)

    mesh = SimpleMesh(nx, ny, Lx * 1e-9, Ly * 1e-9)
    sc = SelfConsistentSolver(mesh, lambda x, y: 12.9 + 0 * x, lambda x, y: 0.067 * M_E + 0 * x,
                              lambda x, y: _v_nm(x * 1e9, y * 1e9) * EV, mixing=0.6)
    sc.eig_tol = 1e-11
    sc.set_solver_parameters(max_iterations=300, tolerance=1e-13)
    out = sc.solve_self_consistent(V_p, V_n, num_quantum_states=k, N_A=NA, N_D=ND)
    assert out["converged"], out["convergence_history"][-5:]
    phi_err = np.abs(out["potential"] - phi_ref).max() / np.abs(phi_ref).max()
    print(f"outer iterations {out['iterations']}, phi vs oracle {phi_err:.2e}", flush=True)
    assert phi_err <= 1e-9

    g = fixed_point_gradients(sc, out["potential"], V_p, V_n, NA, ND, k, g_phi=g_phi, g_levels=g_E)
    print(f"GMRES {g['gmres_iterations']} its, relres {g['gmres_relres']:.1e}, fixed-point residual "
          f"{g['fixed_point_residual']:.1e}, Sternheimer its {g['sternheimer_iterations'][:4]}...", flush=True)
    assert g["fixed_point_residual"] <= 1e-11
    assert np.abs(g["eigenvalues_eV"] - st["E"]).max() <= 1e-9 * np.abs(st["E"]).max()
    for key in ("V_p", "V_n"):
        assert abs(g[key] - ref[key]) <= 1e-6 * abs(ref[key]), (key, g[key], ref[key])
    for key in ("N_A", "N_D"):
        a = g[key].cpu().numpy()
        err = np.abs(a - ref[key]).max() / np.abs(ref[key]).max()
        print(f"{key} rel err {err:.2e}", flush=True)
        assert err <= 1e-6, key
    # v0 at the quadrature points: the two sides order their points differently,
    # so compare directional derivatives along smooth fields
    xq, yq = (t.cpu().numpy() for t in sc.schrodinger.space.quad_points())
    gv = g["v0_q"].cpu().numpy()
    for f in (lambda x, y: np.cos(x / 13.0) * np.sin(y / 7.0), lambda x, y: np.exp(-((x - 30) ** 2 + y ** 2) / 400.0),
              lambda x, y: 1.0 + 0 * x):
        a, b = float((gv * f(xq, yq)).sum()), float((ref["v0_q"] * f(P.xq, P.yq)).sum())
        assert abs(a - b) <= 1e-6 * max(abs(b), 1e-3 * np.abs(ref["v0_q"]).sum()), (a, b)


def test_sc_fixed_point_autograd_gate_parameters():
    """torch.autograd through the fixed point: gate bumps added to the base
    potential (parameters: amplitude and centre of each), scalar doping and a
    contact voltage, against the oracle's reverse-mode gradients chained with the
    closed-form dV/dtheta."""
    import torch
    from oracle import adjoint_ref, sc_adjoint_ref
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.sc_adjoint import self_consistent_fixed_point
    from qdsim_amd.self_consistent import SelfConsistentSolver
    nx, ny, Lx, Ly, k = 17, 11, 80.0, 50.0, 4
    V_p, V_n, NA, ND = -0.05, 0.2, 2e22, 3e22
    theta = np.array([-0.08, 30.0, 20.0, 0.05, 55.0, 30.0])
    sigma = np.array([8.0, 6.0])
    targets = np.array([-0.2, -0.17, -0.15, -0.14])
    print("SC fixed point autograd: 2 gates, scalar doping, P1 16x10 cells", flush=True)
    # oracle
    P = sc_adjoint_ref.SCProblem(nx, ny, Lx, Ly, 12.9, 0.067, k)
    v0 = _v_nm(P.xq, P.yq) + adjoint_ref.gate_v(theta, sigma, P.xq, P.yq)
    rng = np.random.default_rng(11)
    c_phi = 0.1 * rng.standard_normal(P.n)
    _, st, _ = P.fixed_point(v0, V_p, V_n, NA, ND)
    ref = P.gradients(st, c_phi, 2.0 * (st["E"] - targets))
    dv = adjoint_ref.gate_dv(theta, sigma, P.xq, P.yq)
    g_theta_ref = np.einsum("eq,peq->p", ref["v0_q"], dv)
    # product: autograd
    mesh = SimpleMesh(nx, ny, Lx * 1e-9, Ly * 1e-9)
    sc = SelfConsistentSolver(mesh, lambda x, y: 12.9 + 0 * x, lambda x, y: 0.067 * M_E + 0 * x,
                              lambda x, y: _v_nm(x * 1e9, y * 1e9) * EV, mixing=0.6)
    sc.eig_tol = 1e-11
    sc.set_solver_parameters(max_iterations=300, tolerance=1e-13)
    dev = sc.v0_q.device
    xq, yq = sc.schrodinger.space.quad_points()
    th = torch.tensor(theta, dtype=torch.float64, device=dev, requires_grad=True)
    nd = torch.tensor(ND, dtype=torch.float64, device=dev, requires_grad=True)
    vn = torch.tensor(V_n, dtype=torch.float64, device=dev, requires_grad=True)
    gates = sum(th[3 * g] * torch.exp(-((xq - th[3 * g + 1]) ** 2 + (yq - th[3 * g + 2]) ** 2) / (2 * sigma[g] ** 2))
                for g in range(2))
    phi, E = self_consistent_fixed_point(sc, sc.v0_q + gates, NA, nd, V_p, vn, k)
    loss = ((E - torch.as_tensor(targets, device=dev)) ** 2).sum() + (phi * torch.as_tensor(c_phi, device=dev)).sum()
    loss.backward()
    assert np.abs(E.detach().cpu().numpy() - st["E"]).max() <= 1e-9 * np.abs(st["E"]).max()
    g_theta = th.grad.cpu().numpy()
    err = np.abs(g_theta - g_theta_ref).max() / np.abs(g_theta_ref).max()
    print("gate gradient", g_theta, "rel err", err, flush=True)
    assert err <= 1e-6
    assert abs(float(nd.grad) - ref["N_D"].sum()) <= 1e-6 * abs(ref["N_D"].sum())
    assert abs(float(vn.grad) - ref["V_n"]) <= 1e-6 * abs(ref["V_n"])


def test_sc_fixed_point_gradient_junction_central_difference():
    """A 200 x 100 nm p-n junction on 128 x 64 P1 cells (8,385 nodes), 8 states:
    linear mixing does not converge here (the charge feedback puts Jacobian
    eigenvalues below -1 / alpha), Anderson mixing does in ~20 iterations; the
    adjoint gradients with respect to the contact voltage and a doping direction
    match central differences of the converged loop."""
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.sc_adjoint import fixed_point_gradients
    from qdsim_amd.self_consistent import SelfConsistentSolver
    mesh = SimpleMesh(129, 65, 200e-9, 100e-9)
    x, _ = mesh.get_nodes()
    left = x < 97e-9
    NA, ND = np.where(left, 3e22, 1e21), np.where(left, 1e21, 3e22)
    k, V_p, V_n = 8, 0.0, 0.3
    print("SC junction 128x64 P1, Anderson(6), k=8", flush=True)

    def v_nm(xx, yy):
        return -0.25 * np.exp(-((xx - 100) ** 2 + (yy - 50) ** 2) / (2 * 10.0 ** 2))

    def run(Vn, NA_=NA):
        sc = SelfConsistentSolver(mesh, lambda a, b: 13.9 + 0 * a, lambda a, b: 0.067 * M_E + 0 * a,
                                  lambda a, b: v_nm(a * 1e9, b * 1e9) * EV, mixing=0.5, anderson=6)
        sc.eig_tol = 1e-11
        sc.set_solver_parameters(max_iterations=100, tolerance=1e-13)
        out = sc.solve_self_consistent(V_p, Vn, num_quantum_states=k, N_A=NA_, N_D=ND)
        assert out["converged"], out["convergence_history"][-3:]
        return sc, out
    sc, out = run(V_n)
    print("outer iterations", out["iterations"], flush=True)
    assert out["iterations"] <= 60
    rng = np.random.default_rng(3)
    g_phi, g_E = rng.standard_normal(mesh.num_nodes) / mesh.num_nodes, rng.standard_normal(k)
    g = fixed_point_gradients(sc, out["potential"], V_p, V_n, NA, ND, k, g_phi=g_phi, g_levels=g_E)
    print(f"GMRES {g['gmres_iterations']} its, relres {g['gmres_relres']:.1e}", flush=True)

    def J(o):
        return g_phi @ o["potential"] + g_E @ o["eigenvalues_eV"]
    h = 1e-5
    fd = (J(run(V_n + h)[1]) - J(run(V_n - h)[1])) / (2 * h)
    print("dJ/dV_n", g["V_n"], fd, flush=True)
    assert abs(fd - g["V_n"]) <= 1e-6 * abs(fd)
    dNA = np.where(left, 3e22, 0.0) * rng.uniform(0.5, 1.5, mesh.num_nodes)
    eps = 1e-5
    fd = (J(run(V_n, NA + eps * dNA)[1]) - J(run(V_n, NA - eps * dNA)[1])) / (2 * eps)
    an = float(g["N_A"].cpu().numpy() @ dNA)
    print("dJ/dN_A . d", an, fd, flush=True)
    assert abs(fd - an) <= 1e-5 * abs(fd)
