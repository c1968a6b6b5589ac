"""Mixed-precision block kernels (X in FP64, corrections in FP32) and the
mixed-precision block iteration, through the C-ABI.

The typed Gram stages every row as FP64 (FP32 values convert exactly), so it must
equal the FP64 Gram of the upcast blocks bit for bit; the typed update sums the
FP32 blocks' products in another order than the FP64 kernel, so it is compared to
FP64 rounding; the residual step's norms are FP64 sums of the FP64 residual.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _blocks(n, b, seed, have_p=True):
    g = torch.Generator(device="cuda").manual_seed(seed)

    def r(dt):
        return (torch.rand((n, b), generator=g, dtype=torch.float64, device="cuda") - 0.5).to(dt)
    S = [r(torch.float64), r(torch.float32)] + ([r(torch.float32)] if have_p else [])
    HS = [r(torch.float64), r(torch.float32)] + ([r(torch.float32)] if have_p else [])
    MS = [r(torch.float64), r(torch.float32)] + ([r(torch.float32)] if have_p else [])
    return S, HS, MS


@pytest.mark.parametrize("n", [1, 37, 5000, 100003])
@pytest.mark.parametrize("b", [8, 24, 32])
@pytest.mark.parametrize("have_p", [False, True])
def test_typed_gram_equals_fp64_gram_of_upcast_blocks(n, b, have_p):
    from qdsim_amd import kernels
    S, HS, MS = _blocks(n, b, 11 + n + b, have_p)

    def up(blocks):
        return [t.double() for t in blocks]
    for sym in (True, False):
        G = kernels.gram_blocks_typed(S, HS, MS, sym=sym)
        ref = kernels.gram_blocks(up(S), up(HS), up(MS), sym=sym)
        assert torch.equal(G, ref), (n, b, sym, (G - ref).abs().max().item())
    # and against a plain torch FP64 product
    Sd = torch.cat(up(S), 1)
    ref_t = Sd.T @ torch.cat(up(HS) + up(MS), 1)
    assert torch.allclose(G, ref_t, rtol=1e-12, atol=1e-12 * max(n, 1))


@pytest.mark.parametrize("n", [1, 63, 64, 4099, 100003])
@pytest.mark.parametrize("b", [8, 24])
@pytest.mark.parametrize("have_p", [False, True])
def test_typed_update_matches_fp64(n, b, have_p):
    from qdsim_amd import kernels
    S, _, _ = _blocks(n, b, 5 + n, have_p)
    m = len(S) * b
    rng = np.random.default_rng(n + b)
    C = torch.from_numpy(rng.standard_normal((m, b))).cuda()
    CP = C.clone()
    CP[:b] = 0.0
    Y1 = torch.full((n, b), float("nan"), dtype=torch.float64, device="cuda")
    Y2 = torch.full((n, b), float("nan"), dtype=torch.float32, device="cuda")
    kernels.block_update_typed(S, C, Y1, CP, Y2, c2_first=b)
    Sd = torch.cat([t.double() for t in S], 1)
    ref1 = Sd @ C
    ref2 = Sd @ CP
    # the FP64 block (X) contributes with FP64 products and sums; the FP32 blocks
    # (the corrections) on the FP32 MFMA with C rounded to FP32: each part is
    # bounded by its own precision, relative to its own absolute sum
    scale64 = Sd[:, :b].abs() @ C[:b].abs()
    scale32 = Sd[:, b:].abs() @ C[b:].abs()
    bound = 1e-14 * scale64 + 2e-6 * scale32 + 1e-300
    assert torch.all((Y1 - ref1).abs() <= bound), ((Y1 - ref1).abs() / bound).max().item()
    # Y2 has no FP64 block rows (CP[:b] = 0)
    assert torch.all((Y2.double() - ref2).abs() <= 2e-6 * (Sd.abs() @ CP.abs()) + 1e-300)
    # one output only
    Y1b = torch.empty_like(Y1)
    kernels.block_update_typed(S, C, Y1b)
    assert torch.equal(Y1b, Y1)
    # an all-FP64 basis stays exact to FP64
    S64 = [t.double() for t in S]
    Y1c = torch.empty_like(Y1)
    kernels.block_update_typed(S64, C, Y1c)
    assert torch.all((Y1c - ref1).abs() <= 1e-14 * (Sd.abs() @ C.abs()) + 1e-300)


@pytest.mark.parametrize("n", [1, 777, 100003])
def test_residual_to_fp32_matches_fp64_residual(n):
    from qdsim_amd import kernels
    b = 24
    g = torch.Generator(device="cuda").manual_seed(n)
    HX = torch.rand((n, b), generator=g, dtype=torch.float64, device="cuda")
    MX = torch.rand((n, b), generator=g, dtype=torch.float64, device="cuda")
    lam = torch.rand(b, generator=g, dtype=torch.float64, device="cuda")
    cs = (torch.arange(b, device="cuda") % 3 != 0).double()
    R64 = torch.empty_like(HX)
    R32 = torch.empty((n, b), dtype=torch.float32, device="cuda")
    rn64, mn64 = kernels.residual_norms(HX, MX, lam, R64, cs)
    rn64, mn64 = rn64.clone(), mn64.clone()
    rn32, mn32 = kernels.residual_norms(HX, MX, lam, R32, cs)
    assert torch.equal(rn32, rn64) and torch.equal(mn32, mn64)
    assert torch.equal(R32, R64.float())


def _dot_problem(cells, order, L=40.0):
    from qdsim_amd.mesh import SimpleMesh
    from qdsim_amd.schrodinger import SchrodingerProblem
    from qdsim_amd.units import EV_TO_J, M_E

    def v_nm(x, y):
        return -0.3 * np.exp(-((x - L / 2) ** 2 + (y - L / 2) ** 2) / (2 * 5.0 ** 2))

    def m_rel(x, y):
        return np.where((x - L / 2) ** 2 + (y - L / 2) ** 2 < 8.0 ** 2, 0.041, 0.067)

    mesh = SimpleMesh(cells + 1, cells + 1, L * 1e-9, L * 1e-9, order)
    prob = SchrodingerProblem(mesh, lambda x, y: m_rel(x * 1e9, y * 1e9) * M_E,
                              lambda x, y: v_nm(x * 1e9, y * 1e9) * EV_TO_J)
    return prob, m_rel, v_nm


@pytest.mark.parametrize("order,cells", [(2, 24), (3, 16)])
def test_mixed_scheme_matches_fp64_oracle(order, cells):
    from oracle import eig_ref
    L = 40.0
    prob, m_rel, v_nm = _dot_problem(cells, order, L)
    k = 6
    prob.scheme = "mixed"
    lam, vecs, info = prob.solve(k, mode="mixed")
    assert info["converged"] and "mixed_iters" in info["timings"]
    ref = eig_ref.solve_closed(cells + 1, cells + 1, L, L, order, m_rel, v_nm, k)
    err = np.abs(lam - ref["vals"]) / np.abs(ref["vals"])
    assert err.max() < 1e-9, err
    # same answer as the two-phase scheme
    prob.scheme = "two-phase"
    lam2, _, info2 = prob.solve(k, mode="mixed")
    assert np.abs(lam - lam2).max() <= 1e-9 * np.abs(lam2).max()
