"""Host Rayleigh-Ritz (qds_host_rayleigh_ritz, C++ in the library) against LAPACK.

The projected pencil of a block iteration, G = [S^T H S | S^T M S], has exactly
the generalized eigenpairs scipy.linalg.eigh(GH, GM) returns; the host solver must
give the same lowest b eigenvalues and an M-orthonormal C that diagonalises GH
(eigenvectors of degenerate eigenvalues only up to a rotation).  No GPU needed.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from qdsim_amd import kernels
from qdsim_amd.lobpcg import _chol_rr


def _pencil(m, seed, spectrum=None, cond=1e3, n=400):
    rng = np.random.default_rng(seed)
    if spectrum is None:
        spectrum = np.sort(rng.uniform(-1.0, 5.0, n))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    H = (Q * spectrum) @ Q.T
    Mq, _ = np.linalg.qr(rng.standard_normal((n, n)))
    M = (Mq * np.geomspace(1.0, 10.0, n)) @ Mq.T
    S = rng.standard_normal((n, m)) @ np.diag(np.geomspace(1.0, cond ** -0.5, m))
    S *= rng.uniform(0.1, 10.0, m)[None, :]
    return np.hstack([S.T @ H @ S, S.T @ M @ S])


def _check(G, m, b, theta, C, tol=1e-10):
    GH, GM = G[:, :m], G[:, m:]
    ref = sla.eigh(0.5 * (GH + GH.T), 0.5 * (GM + GM.T), eigvals_only=True)[:b]
    scale = np.abs(ref).max()
    assert np.abs(theta - ref).max() <= tol * scale, np.abs(theta - ref).max() / scale
    assert np.abs(C.T @ GM @ C - np.eye(b)).max() < 1e-9
    assert np.abs(C.T @ GH @ C - np.diag(theta)).max() < 1e-9 * scale


@pytest.mark.parametrize("m,b", [(12, 4), (24, 8), (48, 16), (72, 24), (96, 32), (64, 64)])
@pytest.mark.parametrize("seed", [0, 1])
def test_host_rr_matches_lapack(m, b, seed):
    G = _pencil(m, seed)
    out = kernels.host_rayleigh_ritz(G, m, b, 1e-10)
    assert out is not None
    theta, C = out
    _check(G, m, b, theta, C)
    ref = _chol_rr(G, m, b, 1e-10)
    assert np.abs(theta - ref[0]).max() <= 1e-10 * np.abs(ref[0]).max()


def test_host_rr_degenerate_clusters():
    # exactly degenerate pairs and a triple among the lowest values
    n = 300
    spec = np.sort(np.r_[[-2.0, -1.0, -1.0, 0.5, 0.5, 0.5, 1.0, 1.0 + 1e-12], np.linspace(2.0, 9.0, n - 8)])
    m, b = 48, 16
    G = _pencil(m, 3, spectrum=spec, cond=10.0, n=n)
    # the projection of a generic basis is not degenerate; make it so by taking S
    # inside an invariant subspace: rebuild with S spanning eigenvectors
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    H = (Q * spec) @ Q.T
    S = Q[:, :m] @ np.linalg.qr(rng.standard_normal((m, m)))[0]
    G = np.hstack([S.T @ H @ S, S.T @ S])
    theta, C = kernels.host_rayleigh_ritz(G, m, b, 1e-10)
    _check(G, m, b, theta, C)
    assert np.allclose(theta[:8], spec[:8], atol=1e-12)


def test_host_rr_vanished_columns():
    m, b = 72, 24
    G = _pencil(m, 5)
    for i in (30, 31, 50):   # locked residual columns: zero rows and columns
        G[i, :] = 0.0
        G[:, i] = 0.0
        G[:, m + i] = 0.0
    theta, C = kernels.host_rayleigh_ritz(G, m, b, 1e-10)
    assert np.all(C[[30, 31, 50]] == 0.0)
    keep = np.setdiff1d(np.arange(m), [30, 31, 50])
    Gk = np.hstack([G[np.ix_(keep, keep)], G[np.ix_(keep, m + keep)]])
    _check(Gk, len(keep), b, theta, C[keep])


def test_host_rr_declines_near_dependent_basis():
    m, b = 48, 16
    rng = np.random.default_rng(7)
    n = 300
    S = rng.standard_normal((n, m))
    S[:, 40] = S[:, 3] + 1e-9 * rng.standard_normal(n)   # nearly dependent column
    H = np.diag(np.linspace(1.0, 2.0, n))
    G = np.hstack([S.T @ H @ S, S.T @ S])
    assert kernels.host_rayleigh_ritz(G, m, b, 1e-10) is None
    assert _chol_rr(G, m, b, 1e-10) is None   # LAPACK's dpocon declines it as well


def test_host_rr_is_used_by_the_block_iteration():
    from qdsim_amd import lobpcg
    G = _pencil(72, 9)
    theta, C, dropped = lobpcg._rayleigh_ritz(G, 72, 24, 1e-10)
    theta2, C2 = kernels.host_rayleigh_ritz(G, 72, 24, 1e-10)
    assert dropped == 0 and np.array_equal(theta, theta2) and np.array_equal(C, C2)
