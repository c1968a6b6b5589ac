"""Device Rayleigh-Ritz (qds_rayleigh_ritz) against the host LAPACK Rayleigh-Ritz of
the block eigensolver, on Gram pairs of random bases: well conditioned (Cholesky),
with vanished columns (locked residuals, an empty P block), nearly dependent (SVQB
with dropped directions), a batch in one launch, and the residual test / freeze."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _gram(case, n, b, rng):
    m = 3 * b
    S = rng.standard_normal((n, m))
    if case == "vanish":
        S[:, b:b + 5] = 0.0
        S[:, 2 * b:] = 0.0
    if case == "illcond":
        S[:, 2 * b] = S[:, 0] + 1e-6 * rng.standard_normal(n)
    h = np.linspace(-1.0, 5.0, n)
    GH = S.T @ (h[:, None] * S)
    GM = S.T @ S
    return np.hstack([GH, GM]), m


def _check(G, theta, C, b, host):
    m = G.shape[0]
    GH, GM = 0.5 * (G[:, :m] + G[:, :m].T), 0.5 * (G[:, m:] + G[:, m:].T)
    th_h, C_h, dropped_h = host
    scale = np.abs(th_h).max()
    assert np.abs(theta - th_h).max() <= 1e-10 * scale, (theta, th_h)
    assert np.abs(C.T @ GM @ C - np.eye(b)).max() <= 1e-9
    assert np.abs(C.T @ GH @ C - np.diag(theta)).max() <= 1e-9 * scale * max(1.0, np.abs(GH).max() / scale)


@pytest.mark.parametrize("case,b", [("good", 24), ("vanish", 24), ("illcond", 24), ("good", 32), ("illcond", 16)])
def test_device_rr_matches_host(case, b):
    from qdsim_amd import kernels
    from qdsim_amd.lobpcg import _chol_rr, _svqb_rr
    rng = np.random.default_rng(7)
    n = 2000
    G, m = _gram(case, n, b, rng)
    drop_tol = 1e-7
    host = _chol_rr(G, m, b, drop_tol) or _svqb_rr(G, m, b, drop_tol)
    print(case, "m", m, "b", b, "host dropped", host[2], flush=True)
    Gd = torch.from_numpy(G.copy()).cuda().reshape(1, m, 2 * m).contiguous()
    lam = torch.zeros((1, b), dtype=torch.float64, device="cuda")
    C = torch.zeros((1, m, b), dtype=torch.float64, device="cuda")
    CP = torch.full((1, m, b), np.nan, dtype=torch.float64, device="cuda")
    st = torch.zeros((1, kernels.rr_status_len(b)), dtype=torch.float64, device="cuda")
    kernels.rayleigh_ritz(Gd, lam, C, st, k=b, tol=1e-6, drop_tol=drop_tol, CP=CP)
    torch.cuda.synchronize()
    theta, Cn, s = lam[0].cpu().numpy(), C[0].cpu().numpy(), st[0].cpu().numpy()
    print("status", s[:6], flush=True)
    _check(G, theta, Cn, b, host)
    assert int(s[2]) == host[2]                      # dropped directions
    assert int(s[3]) == (1 if case == "illcond" else 0)
    cp = CP[0].cpu().numpy()
    if host[2] > 0:
        assert np.all(cp == 0.0)
    else:
        assert np.all(cp[:b] == 0.0) and np.array_equal(cp[b:], Cn[b:])


def test_device_rr_batch_and_residual_test():
    """Three problems in one launch give what each gives alone; the residual test,
    locking mask and freeze follow the solver's formulas."""
    from qdsim_amd import kernels
    from qdsim_amd.lobpcg import _chol_rr
    rng = np.random.default_rng(11)
    b, n, k, tol = 24, 1500, 16, 1e-6
    Gs = [_gram("good", n, b, rng)[0] for _ in range(3)]
    m = 3 * b
    lam0 = np.sort(rng.uniform(-0.3, 0.1, (3, b)), axis=1)
    norms = np.abs(rng.standard_normal((3, 2 * b)))
    norms[:, :b] *= 1e-20                            # tiny residuals ...
    norms[1, :b] *= 1e12                             # ... except problem 1: not converged
    norms[2, 5] = 1e-6                               # problem 2: column 5 above tol
    G = torch.from_numpy(np.stack(Gs)).cuda().contiguous()
    lam = torch.from_numpy(lam0.copy()).cuda().contiguous()
    C = torch.zeros((3, m, b), dtype=torch.float64, device="cuda")
    cs = torch.zeros((3, b), dtype=torch.float64, device="cuda")
    st = torch.zeros((3, kernels.rr_status_len(b)), dtype=torch.float64, device="cuda")
    nd = torch.from_numpy(norms).cuda().contiguous()
    kernels.rayleigh_ritz(G, lam, C, st, k=k, tol=tol, drop_tol=1e-7, norms=nd, colscale=cs, freeze=True)
    torch.cuda.synchronize()
    s, L, Cn, mask = st.cpu().numpy(), lam.cpu().numpy(), C.cpu().numpy(), cs.cpu().numpy()
    for p in range(3):
        sc = max(np.abs(lam0[p, :k]).max(), 1e-300)
        res = np.sqrt(norms[p, :b]) / (np.maximum(np.abs(lam0[p]), sc) * np.sqrt(norms[p, b:]))
        assert np.allclose(s[p, 8:], res, rtol=1e-13)
        assert s[p, 0] == pytest.approx(res[:k].max(), rel=1e-13)
        assert np.array_equal(mask[p], (res >= tol).astype(float))
        conv = res[:k].max() < tol
        assert bool(s[p, 1]) == conv
        if conv:
            # frozen: X kept, Ritz values unchanged
            assert s[p, 3] == 2.0
            assert np.array_equal(Cn[p], np.eye(m, b)) and np.array_equal(L[p], lam0[p])
        else:
            th_h, C_h, _ = _chol_rr(Gs[p], m, b, 1e-7)
            _check(Gs[p], L[p], Cn[p], b, (th_h, C_h, 0))
    assert [bool(x) for x in s[:, 1]] == [True, False, False]
