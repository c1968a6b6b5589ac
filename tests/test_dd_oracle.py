"""Pins the drift-diffusion oracle (oracle/dd_ref.py) against closed forms and
invariants of the reference's discretisation (full_poisson_dd_solver.cpp).

Each check is built so that a plausible slip fails it: the Poisson parabola fixes
the sign, the 1/6-weight factor, eps0 and the Dirichlet rows; the divergence
identity fixes the drift sign and which index carries the gradient; the continuity
parabola fixes the diffusion coefficient mu kT and the load weight A/3."""
import warnings

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import dd_ref, fem_ref, mesh_ref

LX, LY = 200.0, 100.0


def _ref(nx=21, ny=11, doping=0.0, **kw):
    x, y, el = mesh_ref.simple_mesh_p1(nx, ny, LX, LY)
    N = np.full(len(x), doping) if np.isscalar(doping) else doping(x, y)
    return dd_ref.DDRef(x, y, el, LX, N, **kw)


def test_contacts_follow_the_tolerance_rule():
    d = _ref()
    assert np.array_equal(d.left, np.nonzero(d.x == 0.0)[0])
    assert np.array_equal(d.right, np.nonzero(d.x == LX)[0])


def test_poisson_step_uniform_charge_parabola():
    """K phi = -(1/2) M rho: -eps phi'' = -rho / 2 with phi(0) = V_p, phi(L) = V_n;
    the P1 solution on the structured mesh is nodally exact."""
    d = _ref(doping=0.0)
    N0 = 3.0e4
    d.n[:] = 1e5
    d.p[:] = 1e5
    d.doping[:] = N0                          # rho = q N0
    V_p, V_n = 0.2, -0.1
    d.poisson_step(V_p, V_n)
    eps = 12.9 * dd_ref.EPS0_F_PER_CM
    rho = dd_ref.Q * N0
    exact = V_p + (V_n - V_p) * d.x / LX + rho / (4.0 * eps) * d.x * (d.x - LX)
    scale = np.abs(exact).max()
    assert np.abs(d.phi - exact).max() <= 1e-10 * scale
    # the quadratic term is not negligible: the check sees the charge
    assert np.abs(rho / (4.0 * eps) * (LX / 2) ** 2) > 10 * (abs(V_p) + abs(V_n))


def test_poisson_step_zero_charge_is_linear_ramp():
    d = _ref(doping=0.0)
    d.n[:] = 7e5
    d.p[:] = 7e5
    d.poisson_step(0.3, 0.8)
    assert np.abs(d.phi - (0.3 + 0.5 * d.x / LX)).max() <= 1e-12


def test_field_and_currents_of_linear_fields():
    d = _ref()
    b, c = 2.5e-3, -1.5e-3
    d.phi = 0.1 + b * d.x + c * d.y
    d.electric_field()
    assert np.allclose(d.E, [-b, -c], rtol=1e-12, atol=0)
    d.phi_n = -0.2 + 4e-3 * d.x - 1e-3 * d.y
    d.phi_p = 0.05 - 2e-3 * d.x
    d.n[:] = 2e16
    d.p[:] = 3e15
    d.currents()
    mu_n, mu_p = 8500.0, 400.0
    assert np.allclose(d.Jn, dd_ref.Q * mu_n * 2e16 * np.array([4e-3, -1e-3]), rtol=1e-12, atol=0)
    assert np.allclose(d.Jp, dd_ref.Q * mu_p * 3e15 * np.array([-2e-3, 0.0]), rtol=1e-12, atol=1e-30)


def test_boltzmann_statistics_return_the_initial_carriers():
    """With phi = 0, Boltzmann statistics of the initial quasi-Fermi levels give back
    the mass-action carriers (where the floor does not act)."""
    N = lambda x, y: np.where(x < LX / 2, -1e17, 2e16)
    d = _ref(doping=N)
    n0, p0 = d.n.copy(), d.p.copy()
    ni = d.ni()
    # mass action where neither carrier is floored
    ok = (n0 > 1e5) & (p0 > 1e5)
    assert np.allclose(n0[ok] * p0[ok], ni[ok] ** 2, rtol=1e-12)
    d.phi[:] = 0.0
    d.carriers()
    assert np.allclose(d.n, n0, rtol=1e-12) and np.allclose(d.p, p0, rtol=1e-12)


def test_fermi_dirac_tends_to_boltzmann():
    eta = np.array([-40.0, -30.0, -20.0])
    assert np.allclose(dd_ref.fermi_dirac_half(eta) / np.exp(eta), 1.0, rtol=1e-8)
    # degenerate side: eta^1.5 / (1.5 + 0.27 eta^1.5) -> 1 / 0.27
    assert dd_ref.fermi_dirac_half(np.array([1e12]))[0] == pytest.approx(1 / 0.27, rel=1e-6)


def test_std_max_keeps_nan():
    v = dd_ref.std_max(np.array([np.nan, 1.0, 1e9]), 1e5)
    assert np.isnan(v[0]) and v[1] == 1e5 and v[2] == 1e9


def test_continuity_operator_invariants():
    d = _ref()
    n = d.n_nodes
    mu_n = np.full(n, 8500.0)
    mu_p = np.full(n, 400.0)
    G = np.full(n, 2.0)
    Ex, Ey = 3e-3, -1e-3
    d.E = np.tile([Ex, Ey], (n, 1))
    Kn, Kp, f = d.continuity_matrices(mu_n, mu_p, G)
    # constants are in the kernel of diffusion and drift
    assert np.abs(Kn @ np.ones(n)).max() <= 1e-9 * abs(Kn).max()
    assert np.abs(Kp @ np.ones(n)).max() <= 1e-9 * abs(Kp).max()
    # E = 0: mu kT times the P1 stiffness of an independent assembler
    d.E = np.zeros((n, 2))
    Kn0, Kp0, _ = d.continuity_matrices(mu_n, mu_p, G)
    vx, vy = d.x[d.conn], d.y[d.conn]
    K1, _ = fem_ref.p1_centroid_matrices(vx, vy, np.ones(len(d.conn)), np.ones(len(d.conn)))
    S = fem_ref.assemble(d.conn, n, K1)
    assert abs(Kn0 - 8500.0 * dd_ref.KT * S).max() <= 1e-12 * abs(Kn0).max()
    assert abs(Kp0 - 400.0 * dd_ref.KT * S).max() <= 1e-12 * abs(Kp0).max()
    # divergence identity: 1^T K_drift u = +/- mu E . int grad u = +/- mu E_x |Omega| for u = x
    area = LX * LY
    u = d.x
    dn = np.ones(n) @ ((Kn - Kn0) @ u)
    dp = np.ones(n) @ ((Kp - Kp0) @ u)
    assert dn == pytest.approx(8500.0 * Ex * area, rel=1e-10)
    assert dp == pytest.approx(-400.0 * Ex * area, rel=1e-10)
    v = d.y
    assert np.ones(n) @ ((Kn - Kn0) @ v) == pytest.approx(8500.0 * Ey * area, rel=1e-10)
    # load: sum of G_c A / 3 over the three nodes of every element = G |Omega|
    assert f.sum() == pytest.approx(2.0 * area, rel=1e-12)
    # a linear G integrates exactly under the centroid rule
    _, _, f2 = d.continuity_matrices(mu_n, mu_p, 1.0 + 0.01 * d.x)
    assert f2.sum() == pytest.approx(area * (1.0 + 0.01 * LX / 2), rel=1e-12)


def test_continuity_step_parabola():
    """E = 0 and uniform G: -D phi'' = G with the contact values at x = 0, L, so
    phi = linear + G / (2 D) x (L - x), nodally exact."""
    G0 = 4.0
    d = _ref(doping=0.0, g_r=lambda x, y, n, p, m: G0)
    V_p, V_n = 0.0, 0.25
    d.E = np.zeros((d.n_nodes, 2))
    d.drift_diffusion_step(V_p, V_n)
    ni = d.ni()[0]
    left_n = V_p + dd_ref.KT * np.log(ni * ni / 1e17 / 4.7e17)
    right_n = V_n
    Dn = 8500.0 * dd_ref.KT
    exact_n = left_n + (right_n - left_n) * d.x / LX + G0 / (2 * Dn) * d.x * (LX - d.x)
    assert np.abs(d.phi_n - exact_n).max() <= 1e-10 * np.abs(exact_n).max()
    left_p = V_p - 1.424
    right_p = V_n - 1.424 - dd_ref.KT * np.log(1e17 / 7.0e18)
    Dp = 400.0 * dd_ref.KT
    exact_p = left_p + (right_p - left_p) * d.x / LX + G0 / (2 * Dp) * d.x * (LX - d.x)
    assert np.abs(d.phi_p - exact_p).max() <= 1e-10 * np.abs(exact_p).max()
    # the load term is not negligible
    assert G0 / (2 * Dn) * (LX / 2) ** 2 > 1.0


def test_ldlt_lower_matrix_against_dense_tril():
    A = sp.random(12, 12, density=0.4, random_state=7, format="csr") + sp.eye(12)
    A = A.tocsr()
    D = A.toarray()
    S = dd_ref.ldlt_lower_matrix(A).toarray()
    expect = np.tril(D) + np.tril(D, -1).T
    assert np.array_equal(S, expect)
    assert np.array_equal(S, S.T)


def test_reference_loop_stops_after_one_iteration():
    """Iteration 0 solves no Poisson equation and leaves phi on the ramp, so the
    reference's |phi - phi_old| / |phi| test stops the loop there."""
    d = _ref(doping=lambda x, y: np.where(x < LX / 2, -1e10, 1e10), g_r=lambda x, y, n, p, m: 0.0)
    it, err = d.solve(0.0, 0.4)
    assert it == 1 and err == 0.0 and d.poisson_solves == 0
    assert np.allclose(d.phi, 0.4 * d.x / LX, rtol=0, atol=1e-15)


def test_damping_and_forced_iterations():
    d = _ref(doping=lambda x, y: np.where(x < LX / 2, -1e10, 1e10), g_r=lambda x, y, n, p, m: 0.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        it, err = d.solve(0.0, 0.4, min_iter=3, max_iter=3)
    assert it == 3 and d.poisson_solves == 2
