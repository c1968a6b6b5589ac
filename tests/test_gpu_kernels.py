"""HIP kernels vs FP64 numpy/scipy references (run through the C-ABI)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu


def _rand_csr(n, density, seed, cplx=False):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format='csr')
    A = A + sp.identity(n, format='csr')
    A.sort_indices()
    if cplx:
        B = A.copy()
        B.data = A.data + 1j * rng.standard_normal(A.nnz)
        return B
    return A


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.complex128])
@pytest.mark.parametrize("b", [1, 4, 8, 12, 32])
@pytest.mark.parametrize("fused", [False, True])
def test_spmm(dtype, b, fused):
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    cplx = dtype.is_complex
    n = 3001
    A = _rand_csr(n, 0.004, 1, cplx)
    B = _rand_csr(n, 0.004, 1, cplx)
    B.data = B.data[: 0] if False else (A.data * 0.5 + 1.0)   # same pattern, other values
    rng = np.random.default_rng(2)
    ld = b + 4 if b % 4 == 0 else b
    Xh = rng.standard_normal((n, ld)) + (1j * rng.standard_normal((n, ld)) if cplx else 0)
    X = torch.from_numpy(Xh).to(dev, dtype)[:, :b]
    ip = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ix = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    va = torch.from_numpy(A.data).to(dev, dtype)
    vb = torch.from_numpy(B.data).to(dev, dtype)
    Ya = torch.zeros((n, ld), dtype=dtype, device=dev)[:, :b]
    Yb = torch.zeros((n, ld), dtype=dtype, device=dev)[:, :b]
    if b == 1:
        X = X[:, 0].contiguous(); Ya = torch.zeros(n, dtype=dtype, device=dev); Yb = torch.zeros_like(Ya)
    kernels.spmm(ip, ix, n, A.nnz, va, X, Ya, vb if fused else None, Yb if fused else None)
    Xr = X.cpu().numpy().astype(np.complex128 if cplx else np.float64)
    ref_a = A @ Xr
    tol = 1e-5 if dtype in (torch.float32, torch.complex64) else 1e-12
    scale = abs(A) @ np.abs(Xr)
    assert np.all(np.abs(Ya.cpu().numpy() - ref_a) <= tol * (scale + 1e-30))
    if fused:
        ref_b = B @ Xr
        scale_b = abs(B) @ np.abs(Xr)
        assert np.all(np.abs(Yb.cpu().numpy() - ref_b) <= tol * (scale_b + 1e-30))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("m", [8, 24, 48, 72, 96])
@pytest.mark.parametrize("sym", [False, True])
def test_gram(dtype, m, sym):
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    n = 5003
    rng = np.random.default_rng(3)
    S = torch.from_numpy(rng.standard_normal((n, m + 8))).to(dev, dtype)[:, :m]
    if sym:
        B1 = S  # A^T A symmetric
        B2 = torch.from_numpy(rng.standard_normal((n, m))).to(dev, dtype)
        B2 = (S * 2.0).contiguous()
    else:
        B1 = torch.from_numpy(rng.standard_normal((n, m + 4))).to(dev, dtype)[:, :m - 4 if m > 8 else m]
        B2 = torch.from_numpy(rng.standard_normal((n, 13))).to(dev, dtype)
    G = kernels.gram(S, B1, B2, sym=sym).cpu().numpy()
    Sd = S.double().cpu().numpy()
    ref = Sd.T @ np.concatenate([B1.double().cpu().numpy(), B2.double().cpu().numpy()], 1)
    assert np.allclose(G, ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("b", [8, 12, 16, 32])
@pytest.mark.parametrize("two", [False, True])
def test_block_update(dtype, b, two):
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    n = 4099
    m = 3 * b
    rng = np.random.default_rng(4)
    S = torch.from_numpy(rng.standard_normal((n, m))).to(dev, dtype)
    C1 = torch.from_numpy(rng.standard_normal((m, b))).to(dev, torch.float64)
    C2 = torch.from_numpy(rng.standard_normal((m, b))).to(dev, torch.float64)
    out = torch.full((n, 3 * b), 7.0, dtype=dtype, device=dev)
    Y1 = out[:, :b]
    Y2 = out[:, 2 * b:]
    kernels.block_update(S, C1, Y1, C2 if two else None, Y2 if two else None)
    Sd = S.double().cpu().numpy()
    C1d = C1.to(dtype).double().cpu().numpy()
    tol = 2e-5 if dtype == torch.float32 else 1e-12
    ref1 = Sd @ C1d
    assert np.abs(Y1.double().cpu().numpy() - ref1).max() <= tol * np.abs(Sd).max() * np.abs(C1d).sum(0).max()
    assert torch.all(out[:, b:2 * b] == 7.0)
    if two:
        ref2 = Sd @ C2.to(dtype).double().cpu().numpy()
        assert np.abs(Y2.double().cpu().numpy() - ref2).max() <= tol * np.abs(Sd).max() * np.abs(ref2).max() * 10
    else:
        assert torch.all(out[:, 2 * b:] == 7.0)


@pytest.mark.parametrize("dtype,b", [(torch.float32, 72), (torch.float64, 56), (torch.float64, 40),
                                     (torch.float64, 64)])
@pytest.mark.parametrize("sym", [False, True])
def test_gram_wide_column_blocks(dtype, b, sym):
    """Operands of more than 9 staging pieces (the complex solver's real views):
    one launch per output column group, each with only the pieces it reads."""
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    n = 3001
    rng = np.random.default_rng(5 + b)
    blk = lambda: [torch.from_numpy(rng.standard_normal((n, b))).to(dev, dtype) for _ in range(3)]
    S = blk()
    H = S if sym else blk()
    M = [2.0 * x for x in S] if sym else blk()       # exact scaling keeps S^T M symmetric
    print(f"wide gram {dtype} 3 x {b} columns sym={sym}", flush=True)
    G = kernels.gram_blocks(S, H, M, sym=sym).cpu().numpy()
    cat = lambda L: torch.cat([x.double() for x in L], 1).cpu().numpy()
    ref = cat(S).T @ np.concatenate([cat(H), cat(M)], 1)
    assert np.allclose(G, ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("b", [8, 24, 56])
@pytest.mark.parametrize("nblk", [1, 2, 3])
def test_gram_and_update_column_blocks(dtype, b, nblk):
    """Column-block forms against FP64 products of the concatenated matrices; the
    blocks have different leading dimensions (one is a strided view)."""
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    n = 6007
    rng = np.random.default_rng(11 + b + nblk)

    def blocks(w_extra):
        out = []
        for j in range(nblk):
            full = torch.from_numpy(rng.standard_normal((n, b + (w_extra if j == 1 else 0)))).to(dev, dtype)
            out.append(full[:, :b])
        return out
    S, H, M = blocks(4), blocks(0), blocks(8)
    cat = lambda L: torch.cat([x.double() for x in L], 1).cpu().numpy()
    Sd, Hd, Md = cat(S), cat(H), cat(M)
    G = kernels.gram_blocks(S, H, M).cpu().numpy()
    ref = Sd.T @ np.concatenate([Hd, Md], 1)
    assert np.allclose(G, ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())
    # symmetric halves: A^T [A | 2A] computed from the upper tiles only
    Gs = kernels.gram_blocks(S, S, [2.0 * x for x in S], sym=True).cpu().numpy()
    refs = Sd.T @ np.concatenate([Sd, 2 * Sd], 1)
    assert np.allclose(Gs, refs, rtol=1e-12, atol=1e-10 * np.abs(refs).max())
    m = nblk * b
    C1 = torch.from_numpy(rng.standard_normal((m, b))).to(dev, torch.float64)
    C2 = torch.from_numpy(rng.standard_normal((m, b))).to(dev, torch.float64)
    Y1 = torch.full((n, b), 7.0, dtype=dtype, device=dev)
    Y2 = torch.full((n, b), 7.0, dtype=dtype, device=dev)
    kernels.block_update_blocks(S, C1, Y1, C2, Y2)
    tol = 2e-5 if dtype == torch.float32 else 1e-12
    for Y, C in ((Y1, C1), (Y2, C2)):
        Cd = C.to(dtype).double().cpu().numpy()
        r = Sd @ Cd
        assert np.abs(Y.double().cpu().numpy() - r).max() <= tol * np.abs(Sd).max() * np.abs(Cd).sum(0).max()
    # leading zero rows of C2 declared to the kernel (the P update of the eigensolver)
    C2z = C2.clone()
    C2z[:b] = 0.0
    kernels.block_update_blocks(S, C1, Y1, C2z, Y2, c2_first=b)
    Cd = C2z.to(dtype).double().cpu().numpy()
    r = Sd @ Cd
    assert np.abs(Y2.double().cpu().numpy() - r).max() <= tol * np.abs(Sd).max() * np.abs(Cd).sum(0).max()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.complex128])
def test_residual_and_coldot(dtype):
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    n, b = 2047, 12
    rng = np.random.default_rng(5)
    cplx = dtype.is_complex
    mk = lambda: rng.standard_normal((n, b)) + (1j * rng.standard_normal((n, b)) if cplx else 0)
    HX = torch.from_numpy(mk()).to(dev, dtype)
    MX = torch.from_numpy(mk()).to(dev, dtype)
    lam = torch.from_numpy(rng.standard_normal(b) + (1j * rng.standard_normal(b) if cplx else 0)).to(
        dev, torch.complex128 if cplx else torch.float64)
    R = torch.empty_like(HX)
    kernels.block_residual(HX, MX, lam, R)
    ref = HX.cpu().numpy().astype(np.complex128) - MX.cpu().numpy().astype(np.complex128) * lam.cpu().numpy()[None]
    tol = 1e-5 if dtype in (torch.float32, torch.complex64) else 1e-13
    assert np.abs(R.cpu().numpy() - ref).max() < tol * 10
    d = kernels.coldot(HX, MX, conj=True).cpu().numpy()
    h = HX.cpu().numpy().astype(np.complex128); mm = MX.cpu().numpy().astype(np.complex128)
    assert np.allclose(d, (np.conj(h) * mm).sum(0) if cplx else (h * mm).sum(0).real, rtol=1e-10, atol=1e-9)
    if cplx:
        d2 = kernels.coldot(HX, MX, conj=False).cpu().numpy()
        assert np.allclose(d2, (h * mm).sum(0), rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("b", [4, 12, 24, 32])
def test_residual_norms(dtype, b):
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    n = 70001
    rng = np.random.default_rng(7)
    buf = torch.from_numpy(rng.standard_normal((n, 3 * b))).to(dev, dtype)
    HX, MX = buf[:, :b], buf[:, b:2 * b]
    lam = torch.from_numpy(rng.standard_normal(b)).to(dev, torch.float64)
    R = torch.empty((n, b), dtype=dtype, device=dev)
    rn, mn = kernels.residual_norms(HX, MX, lam, R)
    h = HX.double().cpu().numpy(); m = MX.double().cpu().numpy(); l = lam.cpu().numpy()
    ref = h - m * l.astype(np.float32 if dtype == torch.float32 else np.float64)[None]
    tol = 1e-6 if dtype == torch.float32 else 1e-13
    assert np.abs(R.double().cpu().numpy() - ref).max() < tol * 10
    Rd = R.double().cpu().numpy()
    assert np.allclose(rn.cpu().numpy(), (Rd ** 2).sum(0), rtol=1e-10)
    assert np.allclose(mn.cpu().numpy(), (m ** 2).sum(0), rtol=1e-10)


def test_runtime_check_runs_a_probe_kernel():
    """The library's kernels run through the same HIP runtime as torch's buffers (one
    libamdhip64 image mapped in this process)."""
    from qdsim_amd import _lib
    assert _lib.lib.qds_runtime_check() == 0, _lib.lib.qds_last_error()
    assert _lib.lib.qds_runtime_images() == 1
