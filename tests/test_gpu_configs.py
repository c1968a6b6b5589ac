"""The five benchmark configurations at their own workloads (or, where the CPU
oracle cannot follow, at a reduced cell count of the same physical problem plus
size-independent checks at full size)."""
import numpy as np
import pytest
import torch

from oracle import adjoint_ref, cap_ref, eig_ref

pytestmark = pytest.mark.gpu

EV = 1.602176634e-19
M_E = 9.1093837015e-31
HBAR = 1.054571817e-34


def _problem(nx, ny, Lx, Ly, order, mfun, vfun):
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.schrodinger import SchrodingerProblem
    mesh = SimpleMesh(nx, ny, Lx * 1e-9, Ly * 1e-9, order)
    return SchrodingerProblem(mesh, lambda x, y: mfun(x * 1e9, y * 1e9) * M_E,
                              lambda x, y: vfun(x * 1e9, y * 1e9) * EV)


def _subspace_ok(vecs, ref, tol=1e-6):
    C = vecs.T @ (ref["M"] @ ref["vecs"])
    sv = np.linalg.svd(C, compute_uv=False)
    return sv.min() > 1 - tol, sv.min()


# ---- config 1: P1 quantum well, cell-count reading 64 x 48 cells ----------------
@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_config1_cell_reading(mode):
    def well(x, y):
        return np.where((x > 5.0) & (x < 15.0), -0.06, 0.0) + 0 * y

    def mass(x, y):
        return 0.067 + 0 * x
    print(f"config 1, 64 x 48 cells (65 x 49 nodes), P1, {mode}", flush=True)
    prob = _problem(65, 49, 25.0, 20.0, 1, mass, well)
    ref = eig_ref.solve_closed(65, 49, 25.0, 20.0, 1, mass, well, 5)
    assert len(ref["x"]) == 3185
    lam, vecs, info = prob.solve(5, mode=mode)
    err = np.abs(lam - ref["vals"]) / np.abs(ref["vals"])
    print("rel err", err, flush=True)
    assert err.max() <= 1e-9 and info["converged"]
    ok, smin = _subspace_ok(vecs.cpu().numpy(), ref)
    assert ok, smin


# ---- config 2: P3 Gaussian dot, 100 x 100 nm, 16 states ------------------------
def _dot_v(x, y):
    return -0.3 * np.exp(-((x - 50.0) ** 2 + (y - 50.0) ** 2) / (2 * 5.0 ** 2))


def _dot_m(x, y):
    return np.where((x - 50.0) ** 2 + (y - 50.0) ** 2 < 8.0 ** 2, 0.041, 0.067)


def test_config2_p3_dot_against_oracle_96():
    """The config-2 physics (100 nm, sharp-interface mass, 16 states) at 96^2 P3
    cells (82,369 free DOF), where the FP64 shift-invert oracle runs in seconds."""
    print("config 2 at 96^2 cells, P3, k=16", flush=True)
    prob = _problem(97, 97, 100.0, 100.0, 3, _dot_m, _dot_v)
    ref = eig_ref.solve_closed(97, 97, 100.0, 100.0, 3, _dot_m, _dot_v, 16)
    lam, vecs, info = prob.solve(16)
    err = np.abs(lam - ref["vals"]) / np.abs(ref["vals"])
    print("per-eigenvalue rel err", err, "iters", info["timings"], flush=True)
    assert err.max() <= 1e-9 and info["converged"]
    ok, smin = _subspace_ok(vecs.cpu().numpy(), ref)
    assert ok, smin


def test_config2_full_size_residuals_and_fp64_agreement():
    """512^2 P3 cells (2,356,225 free DOF): the mixed solve's pairs have FP64
    residual below 1e-6 under the CSR form of the operator (a product independent of
    the lattice kernel the solver used), and agree with the all-FP64 solve to 1e-9."""
    print("config 2 full size 512^2 P3", flush=True)
    prob = _problem(513, 513, 100.0, 100.0, 3, _dot_m, _dot_v)
    lam, vecs, info = prob.solve(16)
    assert info["converged"]
    op = prob.operator(torch.float64, lattice=False)
    X = vecs[prob.old_of_new].contiguous()
    HX, MX = torch.empty_like(X), torch.empty_like(X)
    op.apply(X, HX, MX)
    lam_d = torch.as_tensor(lam, device=X.device)
    R = HX - MX * lam_d[None]
    s = np.abs(lam).max()
    res = (R.norm(dim=0) / (s * MX.norm(dim=0))).cpu().numpy()
    print("residuals", res.max(), flush=True)
    assert res.max() < 1e-6
    # M-orthonormal
    G = (X.T @ MX).cpu().numpy()
    assert np.abs(G - np.eye(16)).max() < 1e-8
    lam64, _, info64 = prob.solve(16, mode="fp64")
    assert info64["converged"]
    err = np.abs(lam - lam64) / np.abs(lam64)
    print("mixed vs fp64", err.max(), flush=True)
    assert err.max() <= 1e-9


# ---- config 3: CAP resonances near 0.05 eV -------------------------------------
def _dbw_v(x, y):
    """Double barrier along x: 0.2 eV barriers, 3 nm wide, inner faces 10 nm apart,
    centred in the 80 nm domain."""
    d = np.abs(x - 40.0)
    return np.where((d >= 5.0) & (d <= 8.0), 0.2, 0.0) + 0 * y


def _cap(Lx, Ly, eta=0.05, width=10.0):
    def f(x, y):
        dx = np.maximum(0.0, np.maximum(width - x, x - (Lx - width)))
        dy = np.maximum(0.0, np.maximum(width - y, y - (Ly - width)))
        return eta * (np.maximum(dx, dy) / width) ** 2
    return f


def _cap_problem(cx, cy):
    prob = _problem(cx + 1, cy + 1, 80.0, 60.0, 2, lambda x, y: 0.067 + 0 * x, _dbw_v)
    cfun = _cap(80.0, 60.0)
    prob.set_cap_ev(prob.space.sample(lambda x, y: cfun(x * 1e9, y * 1e9), "quadrature"))
    prob.assemble()
    return prob, cfun


def test_config3_cap_against_eigs_96x72():
    """20 resonances nearest 0.05 eV on the config-3 device at 96 x 72 P2 cells
    (27,313 free DOF) against scipy eigs(H - iW, 20, M, sigma=0.05), with lifetimes
    hbar / 2|Im E|."""
    from qdsim_amd.cap import solve_cap
    print("config 3 at 96 x 72 cells, P2, 20 states nearest 0.05 eV", flush=True)
    prob, cfun = _cap_problem(96, 72)
    lam, X = solve_cap(prob, 20, target=0.05)
    ref = cap_ref.cap_eigs(97, 73, 80.0, 60.0, 2, lambda x, y: 0.067 + 0 * x, _dbw_v, cfun, 20, 0.05)
    a = np.sort_complex(lam)
    r = np.sort_complex(ref["vals"])
    err = np.abs(a - r) / np.abs(r)
    print("eigenvalues", a, "max rel err", err.max(), flush=True)
    assert err.max() <= 1e-9
    assert np.all(a.imag < 0)
    tau, tau_ref = HBAR / (2 * np.abs(a.imag) * EV), HBAR / (2 * np.abs(r.imag) * EV)
    # |d Im E| <= 1e-9 |E| bounds the lifetime error by 1e-9 |E| / |Im E| (relative)
    bound = 1e-9 * np.abs(r) / np.abs(r.imag)
    print("lifetimes (fs)", tau * 1e15, "rel err", np.abs(tau - tau_ref) / tau_ref, flush=True)
    assert np.all(np.abs(tau - tau_ref) / tau_ref <= 2 * bound)


def test_config3_full_size_residuals():
    """384 x 288 P2 cells (441,025 free DOF): residuals of the 20 returned pairs under
    the FP64 pencil evaluated by scipy on the host, and complex symmetry."""
    from qdsim_amd.cap import solve_cap
    print("config 3 full size 384 x 288 P2", flush=True)
    prob, _ = _cap_problem(384, 288)
    lam, X = solve_cap(prob, 20, target=0.05)
    sp_ = prob.space
    H, W, M = sp_.to_scipy(prob.H), sp_.to_scipy(prob.W), sp_.to_scipy(prob.M)
    free = prob.old_of_new.cpu().numpy()
    A = (H - 1j * W)[free][:, free]
    Mf = M[free][:, free]
    assert abs(A - A.T).max() == 0
    Xf = X[free]
    MX = Mf @ Xf
    R = A @ Xf - MX * lam[None]
    rel = np.linalg.norm(R, axis=0) / (np.abs(lam).max() * np.linalg.norm(MX, axis=0))
    dist = np.sort(np.abs(lam - 0.05))
    print("eigenvalues", np.sort_complex(lam), "residual", rel.max(), flush=True)
    assert rel.max() < 2e-6
    assert np.all(lam.imag < 0) and len(lam) == 20 and dist[-1] < 0.05


# ---- config 4: bias sweep instances of the seeded sweep -------------------------
def test_config4_instances_full_iterations():
    """The smallest instance and a P2 instance of >= 256^2 cells of the seeded
    256-instance sweep (seed 4), 30 outer iterations, 8 states, physical Poisson
    sign: every solve converges, contacts hold the bias, the electron density keeps
    the mean donor density, and the levels are finite and ascending."""
    from qdsim_amd.batch import bias_sweep, bias_sweep_instances
    inst = bias_sweep_instances(256, seed=4)
    small = min(inst, key=lambda s: (s["cells"], s["index"]))
    big = next(s for s in inst if s["cells"] >= 256 and s["order"] == 2)
    print("config 4 instances", [(s["index"], s["cells"], s["order"]) for s in (small, big)], flush=True)
    out = bias_sweep([small, big], num_quantum_states=8, iterations=30, workers=2, keep_fields=True)
    for spec, o in zip((small, big), out):
        assert o["iterations"] == 30 and len(o["convergence_history"]) == 30
        assert o["unconverged_solves"] == 0
        lam = np.asarray(o["eigenvalues_eV"])
        assert len(lam) == 8 and np.all(np.isfinite(lam)) and np.all(np.diff(lam) >= 0)
        from qdsim_amd.mesh import SimpleMesh
        m = SimpleMesh(spec["cells"] + 1, spec["cells"] + 1, spec["Lx"], spec["Ly"], spec["order"])
        x, _ = m.get_nodes()
        phi = o["potential"]
        assert np.all(phi[np.abs(x) < 1e-15] == 0.0)
        assert np.allclose(phi[np.abs(x - spec["Lx"]) < 1e-15], spec["bias"], rtol=0, atol=1e-15)
        n = o["electron_density"]
        left = x < spec["junction"]
        nd_mean = np.mean(np.where(left, 0.0, spec["N_D"]))
        assert abs(np.mean(n) - nd_mean) <= 1e-9 * nd_mean
        print(spec["index"], "final max|dphi|", o["convergence_history"][-1], "breakdowns",
              o["breakdown_solves"], flush=True)


# ---- config 5: gate gradients --------------------------------------------------
def test_config5_gradients_against_oracle_reduced():
    """Parameter sets 0-3 of the seeded config-5 sets (seed 5, 8 gates, 24
    parameters, 8 states) on 100 nm at 48^2 P2 cells: the batched adjoint gradient
    against the FP64 oracle gradient (scipy eigenvectors, oracle quadrature), and a
    few parameters against central differences with step 1e-6."""
    from qdsim_amd.batch import gate_gradients, gate_parameter_sets
    theta, sigma, targets = gate_parameter_sets(64, seed=5)
    assert theta.shape == (64, 24) and len(targets) == 8
    kw = dict(cells=48, order=2, L=100e-9, mode="fp64", tol=1e-10)
    loss, g, E = gate_gradients(theta[:4], sigma[:4], targets, **kw)
    for i in range(4):
        lo, go, Eo = adjoint_ref.loss_and_gradient(48, 2, 100.0, 0.067, theta[i], sigma[i], targets)
        assert np.allclose(E[i], Eo, rtol=1e-9, atol=1e-12)
        gerr = np.abs(g[i] - go).max() / np.abs(go).max()
        print("set", i, "loss", loss[i], "oracle", lo, "grad rel err", gerr, flush=True)
        assert abs(loss[i] - lo) <= 1e-9 * lo
        assert gerr <= 1e-7
    h = 1e-6
    for j in (0, 4, 23):
        tp, tm = theta[1:2].copy(), theta[1:2].copy()
        tp[0, j] += h
        tm[0, j] -= h
        fd = (gate_gradients(tp, sigma[1:2], targets, **kw)[0][0]
              - gate_gradients(tm, sigma[1:2], targets, **kw)[0][0]) / (2 * h)
        print("param", j, "adjoint", g[1, j], "fd", fd, flush=True)
        assert abs(g[1, j] - fd) <= 1e-4 * abs(fd) + 1e-8


def test_config5_full_size_64_sets():
    """All 64 sets at the config-5 size (256^2 P2 cells, 263,169 DOF, mixed precision,
    tol 1e-9): finite gradients for every set, and set 7 against central
    differences on two parameters."""
    from qdsim_amd.batch import gate_gradients, gate_parameter_sets
    theta, sigma, targets = gate_parameter_sets(64, seed=5)
    loss, g, E = gate_gradients(theta, sigma, targets)
    assert g.shape == (64, 24) and np.all(np.isfinite(g)) and np.all(np.isfinite(E))
    assert np.all(np.diff(E, axis=1) >= 0)
    h = 1e-6
    for j in (3, 6):
        tp, tm = theta[7:8].copy(), theta[7:8].copy()
        tp[0, j] += h
        tm[0, j] -= h
        fd = (gate_gradients(tp, sigma[7:8], targets)[0][0] - gate_gradients(tm, sigma[7:8], targets)[0][0]) / (2 * h)
        print("set 7 param", j, "adjoint", g[7, j], "fd", fd, flush=True)
        assert abs(g[7, j] - fd) <= 1e-3 * abs(fd) + 1e-7
