"""Lattice (stencil) form of the free-DOF operator on the GPU: qds_lattice_spmm
against the CSR kernel (bitwise) and against scipy's FP64 product (tolerance)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu


def _setup(p, ncx, ncy, seed=0):
    from qdsim_amd import kernels
    from qdsim_amd.lattice import Lattice
    from qdsim_amd.mesh import SimpleMesh
    m = SimpleMesh(ncx + 1, ncy + 1, 1e-7, 1e-7, p)
    conn = np.ascontiguousarray(m.get_elements(), dtype=np.int32)
    pat = kernels.pattern_build(conn, int(m.num_nodes))
    lat = Lattice(m)
    r_ip, r_ix, _ = kernels.pattern_restrict(pat["indptr"], pat["indices"], lat.new_of_old, lat.n_free)
    st, slot = kernels.lattice_stencil(p, lat.nx_i, lat.ny_i, r_ip, r_ix)
    rng = np.random.default_rng(seed)
    va = rng.standard_normal(len(r_ix))
    vb = rng.uniform(0.5, 1.5, len(r_ix))
    return lat, r_ip, r_ix, st, slot, va, vb


def _device_operator(r_ip, r_ix, slot, va, vb, dtype):
    """CSR and lattice value arrays on the device; the lattice values go through
    qds_csr_take with the slot map, as the solver builds them."""
    from qdsim_amd import kernels
    dev = torch.device("cuda")
    ha = torch.from_numpy(va).to(dev)
    hb = torch.from_numpy(vb).to(dev)
    src = torch.from_numpy(slot.astype(np.int32)).to(dev)
    csr = (torch.from_numpy(r_ip.astype(np.int32)).to(dev), torch.from_numpy(r_ix.astype(np.int32)).to(dev),
           ha.to(dtype), hb.to(dtype))
    lat = (kernels.csr_take(src, ha, dtype=dtype), kernels.csr_take(src, hb, dtype=dtype))
    return csr, lat


@pytest.mark.parametrize("p,ncx,ncy", [(1, 37, 23), (2, 19, 13), (3, 24, 11), (3, 1, 3), (3, 40, 1)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("b,ld", [(1, 1), (3, 3), (8, 8), (24, 24), (24, 29), (26, 26)])
@pytest.mark.parametrize("fused", [True, False])
def test_lattice_spmm_matches_csr(p, ncx, ncy, dtype, b, ld, fused):
    from qdsim_amd import kernels
    lat, r_ip, r_ix, st, slot, va, vb = _setup(p, ncx, ncy)
    n = lat.n_free
    print(f"lattice spmm p={p} cells={ncx}x{ncy} n={n} b={b} ld={ld} {dtype} fused={fused}", flush=True)
    (ip, ix, ca, cb), (la, lb) = _device_operator(r_ip, r_ix, slot, va, vb, dtype)
    dev = ip.device
    rng = np.random.default_rng(7)
    Xw = torch.from_numpy(rng.standard_normal((n, ld))).to(dev, dtype)
    X = Xw[:, :b] if b > 1 else Xw[:, 0].contiguous()

    def out():
        if b == 1:
            return torch.full((n,), float("nan"), dtype=dtype, device=dev)
        return torch.full((n, ld), float("nan"), dtype=dtype, device=dev)[:, :b]

    Ya, Yb, Za, Zb = out(), out(), out(), out()
    kernels.lattice_spmm(st, la, X, Ya, lb if fused else None, Yb if fused else None)
    kernels.spmm(ip, ix, n, len(r_ix), ca, X, Za, cb if fused else None, Zb if fused else None)
    torch.cuda.synchronize()
    A = sp.csr_matrix((va, r_ix, r_ip), shape=(n, n))
    B = sp.csr_matrix((vb, r_ix, r_ip), shape=(n, n))
    Xh = X.double().cpu().numpy()
    tol = 1e-5 if dtype == torch.float32 else 1e-12
    pairs = [(Ya, Za, A)] + ([(Yb, Zb, B)] if fused else [])
    for Y, Z, M in pairs:
        y = Y.cpu().numpy()
        assert np.all(np.isfinite(y)), "every row written"
        ref = M @ Xh
        scale = abs(M) @ np.abs(Xh)
        assert np.all(np.abs(y - ref) <= tol * (scale + 1e-30))
        if b > 1:   # the CSR single-vector kernel sums with lane shuffles; blocks sum in column order
            assert torch.equal(Y, Z), "lattice and CSR products differ"


def test_lattice_operator_in_solver_matches_csr():
    """SchrodingerProblem's default lattice operator applies the same matrices as its
    CSR operator (P3 quantum dot, FP32 and FP64)."""
    import bench
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.schrodinger import SchrodingerProblem
    m_star, potential = bench._funcs()
    prob = SchrodingerProblem(SimpleMesh(49, 41, 100e-9, 100e-9, 3), m_star, potential)
    prob.assemble()
    for dtype in (torch.float32, torch.float64):
        op_l = prob.operator(dtype)
        op_c = prob.operator(dtype, lattice=False)
        assert type(op_l).__name__ == "LatticeOperator"
        X = torch.randn((prob.n_free, 24), dtype=dtype, device="cuda")
        outs = [torch.empty_like(X) for _ in range(4)]
        op_l.apply(X, outs[0], outs[1])
        op_c.apply(X, outs[2], outs[3])
        assert torch.equal(outs[0], outs[2]) and torch.equal(outs[1], outs[3])
    # FP64 lattice values rounded to FP32 (the mixed solver's phase switch) equal
    # the FP32 operator built directly
    op_a = prob.operator(torch.float64).astype(torch.float32)
    op_b = prob.operator(torch.float32)
    assert torch.equal(op_a.H, op_b.H) and torch.equal(op_a.M, op_b.M)


@pytest.mark.parametrize("p,ncx,ncy", [(1, 37, 23), (2, 19, 13), (3, 24, 11), (3, 96, 40)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("fused", [True, False])
def test_every_tuning_candidate_gives_the_same_product(p, ncx, ncy, dtype, fused):
    """All block shapes of the autotuner (values in LDS or from HBM, LDS budgets)
    compute the same sums in the same order: bitwise equal to the CSR product, one
    problem and a batch of three."""
    from qdsim_amd import autotune, kernels
    lat, r_ip, r_ix, st, slot, va, vb = _setup(p, ncx, ncy)
    n = lat.n_free
    (ip, ix, ca, cb), (la, lb) = _device_operator(r_ip, r_ix, slot, va, vb, dtype)
    b = 24 if dtype == torch.float32 else 16
    rng = np.random.default_rng(3)
    X = torch.from_numpy(rng.standard_normal((3, n, b))).cuda().to(dtype).contiguous()
    # per-set CSR references: A_i = s_i A (s = 0.5, 1, 2: exact scalings), B shared
    refs = []
    for i, sc in enumerate((0.5, 1.0, 2.0)):
        ra, rb = torch.empty((n, b), dtype=dtype, device="cuda"), torch.empty((n, b), dtype=dtype, device="cuda")
        kernels.spmm(ip, ix, n, int(ix.numel()), ca * sc, X[i], ra, cb if fused else None, rb if fused else None)
        refs.append((ra, rb))
    ref_a, ref_b = refs[1]
    LA = torch.stack([la * 0.5, la, la * 2.0]).contiguous()
    print(f"tuning candidates p={p} {ncx}x{ncy} {dtype} fused={fused}", flush=True)
    for cand in autotune.CANDIDATES:
        ya = torch.full((3, n, b), float("nan"), dtype=dtype, device="cuda")
        yb = torch.full((3, n, b), float("nan"), dtype=dtype, device="cuda")
        try:
            kernels.lattice_spmm_tuned(st, LA, X, ya, lb if fused else None, yb if fused else None, cand)
        except Exception as e:   # a shape the kernel rejects (too large for LDS) is skipped by the tuner
            print("rejected", cand, e)
            continue
        for i in range(3):
            assert torch.equal(ya[i], refs[i][0]), (cand, i)
            if fused:
                assert torch.equal(yb[i], refs[i][1]), (cand, i)
        y1 = torch.empty((n, b), dtype=dtype, device="cuda")
        kernels.lattice_spmm_tuned(st, la, X[1].contiguous(), y1, None, None, cand)
        assert torch.equal(y1, ref_a)
    ya = torch.empty((n, b), dtype=dtype, device="cuda")
    kernels.lattice_spmm_auto(st, la, X[1].contiguous(), ya, lb if fused else None,
                              torch.empty_like(ya) if fused else None)
    assert torch.equal(ya, ref_a)


@pytest.mark.parametrize("p,ncx,ncy", [(1, 37, 23), (2, 19, 13), (3, 24, 11), (3, 64, 40), (3, 1, 3)])
@pytest.mark.parametrize("dtype,rtype", [(torch.float64, torch.float32), (torch.float64, torch.float64),
                                         (torch.float32, torch.float32)])
@pytest.mark.parametrize("b", [4, 24])
@pytest.mark.parametrize("choice", [(1, 53248), (0, 32768), (1, 24576)])
def test_lattice_spmm_residual_epilogue(p, ncx, ncy, dtype, rtype, b, choice):
    """The fused residual step gives the plain product bit for bit, R bit for bit
    as the residual-norms kernel computes it from the stored products, and the same
    column norms (another summation order)."""
    from qdsim_amd import kernels
    lat, r_ip, r_ix, st, slot, va, vb = _setup(p, ncx, ncy)
    n = lat.n_free
    print(f"lattice residual p={p} cells={ncx}x{ncy} n={n} b={b} {dtype}->{rtype} choice={choice}", flush=True)
    _, (la, lb) = _device_operator(r_ip, r_ix, slot, va, vb, dtype)
    dev = la.device
    rng = np.random.default_rng(11)
    X = torch.from_numpy(rng.standard_normal((n, b))).to(dev, dtype)
    lam = torch.from_numpy(rng.uniform(0.5, 2.0, b)).to(dev)
    cs = torch.from_numpy((np.arange(b) % 3 != 1).astype(np.float64)).to(dev)
    HX, MX = torch.empty_like(X), torch.empty_like(X)
    kernels.lattice_spmm_tuned(st, la, X, HX, lb, MX, choice)
    R_ref = torch.empty((n, b), dtype=rtype, device=dev)
    for colscale in (cs, None):
        rn_ref, mn_ref = (t.clone() for t in kernels.residual_norms(HX, MX, lam, R_ref, colscale))
        HX2 = torch.full_like(X, float("nan"))
        MX2 = torch.full_like(X, float("nan"))
        R = torch.full((n, b), float("nan"), dtype=rtype, device=dev)
        rn, mn = kernels.lattice_spmm_residual(st, la, lb, X, HX2, MX2, lam, R, colscale, choice=choice)
        assert torch.equal(HX2, HX) and torch.equal(MX2, MX)
        assert torch.equal(R, R_ref)
        assert torch.allclose(rn, rn_ref, rtol=1e-12, atol=0) and torch.allclose(mn, mn_ref, rtol=1e-12, atol=0)
