"""GPU eigensolver vs the FP64 shift-invert oracle."""
import numpy as np
import pytest
import torch

from oracle import eig_ref

pytestmark = pytest.mark.gpu


def _well(x, y):          # -60 meV well for 5 < x < 15 nm, x in nm
    return np.where((x > 5.0) & (x < 15.0), -0.06, 0.0) + 0 * y


def _const_mass(x, y):
    return 0.067 + 0 * x


def _dot_v(cx, cy):
    return lambda x, y: -0.3 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 5.0 ** 2))


def _dot_m(cx, cy):
    return lambda x, y: np.where((x - cx) ** 2 + (y - cy) ** 2 < 8.0 ** 2, 0.041, 0.067)


def _problem(nx, ny, Lx, Ly, order, mfun, vfun):
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.schrodinger import SchrodingerProblem
    from qdsim_amd.units import EV_TO_J, M_E
    mesh = SimpleMesh(nx, ny, Lx * 1e-9, Ly * 1e-9, order)
    return SchrodingerProblem(mesh, lambda x, y: mfun(x * 1e9, y * 1e9) * M_E,
                              lambda x, y: vfun(x * 1e9, y * 1e9) * EV_TO_J)


def _check(prob, ref, k, mode, rel=1e-9):
    lam, vecs, info = prob.solve(k, mode=mode)
    err = np.abs(lam - ref["vals"]) / np.abs(ref["vals"])
    print(mode, "eigen rel err", err.max(), "iters", info["iterations"], info["timings"])
    assert err.max() <= rel, err
    # eigenvectors: same subspace (M-inner products), robust to degeneracy and sign
    V = vecs.cpu().numpy()
    Mref = ref["M"]
    C = V.T @ (Mref @ ref["vecs"])
    sv = np.linalg.svd(C, compute_uv=False)
    assert np.all(sv > 1 - 1e-6), sv
    return lam, info


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_p1_quantum_well_64x48(mode):
    # node-count reading of FixedOpenSystemSolver(nx=64, ny=48), 25 x 20 nm, P1
    prob = _problem(64, 48, 25.0, 20.0, 1, _const_mass, _well)
    ref = eig_ref.solve_closed(64, 48, 25.0, 20.0, 1, _const_mass, _well, 5)
    _check(prob, ref, 5, mode)


@pytest.mark.parametrize("order,cells", [(2, 32), (3, 16)])
@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_dot_p2_p3(order, cells, mode):
    L = 40.0
    mf, vf = _dot_m(L / 2, L / 2), _dot_v(L / 2, L / 2)
    prob = _problem(cells + 1, cells + 1, L, L, order, mf, vf)
    ref = eig_ref.solve_closed(cells + 1, cells + 1, L, L, order, mf, vf, 8)
    _check(prob, ref, 8, mode)


def test_nested_start_matches_random_start():
    """The coarse-mesh initial block (nested iteration) converges to the same pairs as
    the seeded random start; P2 256 x 256 cells (263,169 DOF), coarse mesh 64 x 64."""
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.schrodinger import SchrodingerProblem
    L = 100e-9

    def m_star(x, y):
        return 0.067 * 9.1093837015e-31 + 0 * x

    def pot(x, y):
        r2 = (x - L / 2) ** 2 + (y - L / 2) ** 2
        e = np.exp(-r2 / (2 * (8e-9) ** 2)) if isinstance(r2, np.ndarray) else torch.exp(-r2 / (2 * (8e-9) ** 2))
        return -0.25 * 1.602176634e-19 * e
    mesh = SimpleMesh(257, 257, L, L, 2)
    prob = SchrodingerProblem(mesh, m_star, pot)
    lam_n, X_n, info_n = prob.solve(8)
    assert prob._coarse is not None, "nested start not used"
    prob.nested = False
    lam_r, X_r, info_r = prob.solve(8)
    print("nested", info_n["timings"], "random", info_r["timings"])
    assert info_n["converged"] and info_r["converged"]
    assert np.allclose(lam_n, lam_r, rtol=1e-10, atol=0)
