"""Mixing of the self-consistent loop (host logic, no GPU): linear mixing is the
reference's update; Anderson mixing converges on a linear map whose Jacobian has
eigenvalues far below -1, where linear mixing with the same alpha diverges, and
reaches the exact fixed point of a linear map in at most n + 1 steps of history."""
import numpy as np
import torch

from qdsim_amd.self_consistent import mix_step


def _linear_map(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    J = (Q * np.linspace(lo, hi, n)) @ Q.T
    c = rng.standard_normal(n)
    fixed = np.linalg.solve(np.eye(n) - J, c)
    return (lambda x: torch.from_numpy(J @ x.numpy() + c)), fixed


def test_linear_mixing_is_the_reference_update():
    phi = torch.tensor([0.0, 1.0, 2.0], dtype=torch.float64)
    new = torch.tensor([0.0, 3.0, 0.0], dtype=torch.float64)
    out = mix_step(phi, new, [], 0.3, 0)
    assert torch.allclose(out, phi + 0.3 * (new - phi), rtol=0, atol=0)


def test_anderson_converges_where_linear_mixing_diverges():
    F, fixed = _linear_map(40, 0, -12.0, 0.5)
    for depth, expect in ((0, False), (8, True)):
        phi = torch.zeros(40, dtype=torch.float64)
        hist = []
        for _ in range(200):
            phi = mix_step(phi, F(phi), hist, 0.3, depth)
            if not torch.all(torch.isfinite(phi)) or float(phi.abs().max()) > 1e12:
                break
        ok = bool(torch.all(torch.isfinite(phi))) and np.abs(phi.numpy() - fixed).max() < 1e-8 * np.abs(fixed).max()
        assert ok == expect, depth


def test_anderson_exact_on_small_linear_map():
    F, fixed = _linear_map(5, 1, -3.0, 0.8)
    phi = torch.zeros(5, dtype=torch.float64)
    hist = []
    for _ in range(7):
        phi = mix_step(phi, F(phi), hist, 0.5, 10)
    assert np.abs(phi.numpy() - fixed).max() < 1e-9 * np.abs(fixed).max()


def test_anderson_singular_history_falls_back():
    phi = torch.ones(4, dtype=torch.float64)
    hist = []
    a = mix_step(phi, phi + 1.0, hist, 0.5, 3)
    b = mix_step(a, a + 1.0, hist, 0.5, 3)     # identical residuals: dF = 0
    assert torch.all(torch.isfinite(b))
