"""Lattice multigrid kernels vs a numpy stencil and as a solver."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _stencil(d, we, wn, u, nx, ny):
    U = u.reshape(ny, nx, -1)
    D = d.reshape(ny, nx, 1); WE = we.reshape(ny, nx, 1); WN = wn.reshape(ny, nx, 1)
    out = D * U
    out[:, :-1] -= WE[:, :-1] * U[:, 1:]
    out[:, 1:] -= WE[:, :-1] * U[:, :-1]
    out[:-1] -= WN[:-1] * U[1:]
    out[1:] -= WN[:-1] * U[:-1]
    return out.reshape(nx * ny, -1)


def _mg(n_cells=64, b=8, dtype=torch.float64, ny_cells=None, **kw):
    from qdsim_amd.lattice import Lattice, LatticeMG
    from qdsim_amd.mesh import SimpleMesh
    ny_cells = n_cells if ny_cells is None else ny_cells
    m = SimpleMesh(n_cells + 1, ny_cells + 1, 50e-9, 50e-9 * ny_cells / n_cells, 1)
    lat = Lattice(m)
    E = m.num_elements
    dev = torch.device("cuda")
    rng = np.random.default_rng(0)
    c = torch.from_numpy(0.5 + 0.5 * rng.random(E)).to(dev)
    v = torch.from_numpy(0.1 * rng.random(E)).to(dev)
    return lat, LatticeMG(lat, c, v, sigma=-0.01, b=b, dtype=dtype, **kw)


def test_mg_apply_matches_stencil():
    from qdsim_amd import _lib, kernels
    lat, mg = _mg()
    lv = mg.levels_data[0]
    nx, ny = lv["nx"], lv["ny"]
    u = torch.randn((nx * ny, mg.b), dtype=torch.float64, device="cuda")
    out = torch.empty_like(u)
    kernels.mg_apply(_lib.F64, mg._structs[0], mg.b, u, out)
    ref = _stencil(lv["d"].cpu().numpy(), lv["we"].cpu().numpy(), lv["wn"].cpu().numpy(), u.cpu().numpy(), nx, ny)
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nu,coarse,bound", [(2, 40, 0.35), (1, 10, 0.6)])
def test_vcycle_converges(nu, coarse, bound):
    from qdsim_amd import _lib, kernels
    lat, mg = _mg(128, nu_pre=nu, nu_post=nu, coarse_sweeps=coarse)
    assert mg.n_levels >= 4
    lv = mg.levels_data[0]
    n = lv["nx"] * lv["ny"]
    f = torch.randn((n, mg.b), dtype=torch.float64, device="cuda")
    x = torch.zeros_like(f)
    r = f.clone()
    norms = [r.norm().item()]
    w = torch.empty_like(f)
    Ax = torch.empty_like(f)
    for _ in range(8):
        mg(r, w)
        x += w
        kernels.mg_apply(_lib.F64, mg._structs[0], mg.b, x, Ax)
        r = f - Ax
        norms.append(r.norm().item())
    factor = (norms[-1] / norms[0]) ** (1 / 8)
    print("V-cycle convergence factor", factor, norms)
    assert factor < bound


def _np_vcycle(levels, f, nu1, nu2, ncoarse, omega):
    """V(nu1, nu2) cycle with damped Jacobi, bilinear P and R = P^T, in numpy FP64:
    the same schedule as qds_mg_vcycle (zero initial guess on every level)."""
    def A(l, u):
        return _stencil(l["d"], l["we"], l["wn"], u, l["nx"], l["ny"])

    def smooth(l, u, f, n):
        for s in range(n):
            u = omega * f / l["d"][:, None] if u is None else u + omega * (f - A(l, u)) / l["d"][:, None]
        return u

    def restrict(l, lc, r):
        R = r.reshape(l["ny"], l["nx"], -1)
        out = np.zeros((lc["ny"], lc["nx"], R.shape[2]))
        w = (0.5, 1.0, 0.5)
        for a in range(3):
            for c in range(3):
                out += w[a] * w[c] * R[a:a + 2 * lc["ny"]:2, c:c + 2 * lc["nx"]:2]
        return out.reshape(lc["nx"] * lc["ny"], -1)

    def prolong(l, lc, e):
        E = e.reshape(lc["ny"], lc["nx"], -1)
        rows = np.zeros((lc["ny"], l["nx"], E.shape[2]))
        rows[:, 1::2] = E
        rows[:, 2:-1:2] = 0.5 * (E[:, :-1] + E[:, 1:])
        rows[:, 0] = 0.5 * E[:, 0]
        rows[:, -1] = 0.5 * E[:, -1]
        out = np.zeros((l["ny"], l["nx"], E.shape[2]))
        out[1::2] = rows
        out[2:-1:2] = 0.5 * (rows[:-1] + rows[1:])
        out[0] = 0.5 * rows[0]
        out[-1] = 0.5 * rows[-1]
        return out.reshape(l["nx"] * l["ny"], -1)

    def cyc(i, f):
        l = levels[i]
        if i == len(levels) - 1:
            return smooth(l, None, f, ncoarse)
        u = smooth(l, None, f, nu1)
        e = cyc(i + 1, restrict(l, levels[i + 1], f - A(l, u)))
        return smooth(l, u + prolong(l, levels[i + 1], e), f, nu2)
    return cyc(0, f)


@pytest.mark.parametrize("b,dtype,cells", [(8, torch.float64, (128, 128)), (24, torch.float32, (128, 128)),
                                           (6, torch.float64, (128, 128)), (6, torch.float32, (128, 128)),
                                           (24, torch.float32, (288, 160)), (24, torch.float64, (160, 288))])
def test_vcycle_matches_numpy_schedule(b, dtype, cells):
    """One V-cycle (multi-launch fine levels + single-workgroup coarse levels)
    against the same cycle in numpy FP64, on square and rectangular lattices."""
    lat, mg = _mg(cells[0], b=b, dtype=dtype, ny_cells=cells[1])
    levels = [{"nx": lv["nx"], "ny": lv["ny"], "d": lv["d"].double().cpu().numpy(),
               "we": lv["we"].double().cpu().numpy(), "wn": lv["wn"].double().cpu().numpy()}
              for lv in mg.levels_data]
    n = levels[0]["nx"] * levels[0]["ny"]
    rng = np.random.default_rng(5)
    f = rng.standard_normal((n, b))
    fd = torch.from_numpy(f).to("cuda", dtype)
    w = torch.empty_like(fd)
    mg(fd, w)
    ref = _np_vcycle(levels, fd.double().cpu().numpy(), mg.nu_pre, mg.nu_post, mg.coarse_sweeps, mg.omega)
    tol = 1e-11 if dtype == torch.float64 else 2e-5
    err = np.abs(w.double().cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("cells,b", [((128, 128), 24), ((288, 160), 8)])
def test_vcycle_fp64_blocks_on_fp32_hierarchy(cells, b):
    """qds_mg_vcycle_io: an FP32 hierarchy applied to FP64 R, W gives bitwise the
    result of converting R to FP32, cycling, and converting W back (the loads and
    stores convert with the same rounding)."""
    lat, mg = _mg(cells[0], b=b, dtype=torch.float32, ny_cells=cells[1], nu_pre=1, nu_post=1)
    assert mg.takes_fp64
    n = mg.levels_data[0]["nx"] * mg.levels_data[0]["ny"]
    g = torch.Generator(device="cuda").manual_seed(3)
    R = torch.randn((n, b), dtype=torch.float64, device="cuda", generator=g)
    W = torch.full_like(R, float("nan"))
    mg(R, W)
    R32 = R.float()
    W32 = torch.empty_like(R32)
    mg(R32, W32)
    torch.cuda.synchronize()
    assert torch.equal(W, W32.double())


def _np_vcycle_neumann_y(levels, f, ncoarse, omega):
    """V(1,1) on a hierarchy that is Dirichlet in x (coarse j at fine 2j+1) and
    Neumann in y (the boundary rows are unknowns; coarse row j at fine 2j)."""
    def A(l, u):
        return _stencil(l["d"], l["we"], l["wn"], u, l["nx"], l["ny"])

    def smooth(l, u, f, n):
        for s in range(n):
            u = omega * f / l["d"][:, None] if u is None else u + omega * (f - A(l, u)) / l["d"][:, None]
        return u

    def restrict(l, lc, r):
        R = r.reshape(l["ny"], l["nx"], -1)
        out = np.zeros((lc["ny"], lc["nx"], R.shape[2]))
        for cy in range(lc["ny"]):
            for dy, wy in ((-1, 0.5), (0, 1.0), (1, 0.5)):
                y = 2 * cy + dy
                if 0 <= y < l["ny"]:
                    row = R[y]
                    out[cy] += wy * (0.5 * row[0:2 * lc["nx"]:2] + row[1:2 * lc["nx"] + 1:2]
                                     + 0.5 * row[2:2 * lc["nx"] + 2:2])
        return out.reshape(lc["nx"] * lc["ny"], -1)

    def prolong(l, lc, e):
        E = e.reshape(lc["ny"], lc["nx"], -1)
        rows = np.zeros((lc["ny"], l["nx"], E.shape[2]))
        rows[:, 1::2] = E
        rows[:, 2:-1:2] = 0.5 * (E[:, :-1] + E[:, 1:])
        rows[:, 0] = 0.5 * E[:, 0]
        rows[:, -1] = 0.5 * E[:, -1]
        out = np.zeros((l["ny"], l["nx"], E.shape[2]))
        out[0::2] = rows
        out[1::2] = 0.5 * (rows[:-1] + rows[1:])
        return out.reshape(l["nx"] * l["ny"], -1)

    def cyc(i, f):
        l = levels[i]
        if i == len(levels) - 1:
            return smooth(l, None, f, ncoarse)
        u = smooth(l, None, f, 1)
        e = cyc(i + 1, restrict(l, levels[i + 1], f - A(l, u)))
        return smooth(l, u + prolong(l, levels[i + 1], e), f, 1)
    return cyc(0, f)


@pytest.mark.parametrize("order,cells", [(1, (64, 32)), (2, (96, 48)), (2, (200, 100))])
def test_poisson_vcycle_neumann_rows_matches_numpy(order, cells):
    """The Poisson hierarchy (Dirichlet columns, Neumann top and bottom rows whose
    boundary rows are coarse points): one device V-cycle against numpy FP64."""
    from qdsim_amd.lattice import Lattice, PoissonLatticeMG
    from qdsim_amd.mesh import SimpleMesh
    m = SimpleMesh(cells[0] + 1, cells[1] + 1, 200e-9, 100e-9, order)
    rng = np.random.default_rng(1)
    eps = torch.from_numpy(10.0 + 5.0 * rng.random(m.num_elements)).to("cuda")
    mg = PoissonLatticeMG(Lattice(m), eps)
    print("poisson hierarchy", [(lv["nx"], lv["ny"]) for lv in mg.levels_data], flush=True)
    assert mg.n_levels >= 3
    levels = [{"nx": lv["nx"], "ny": lv["ny"], "d": lv["d"].cpu().numpy(), "we": lv["we"].cpu().numpy(),
               "wn": lv["wn"].cpu().numpy()} for lv in mg.levels_data]
    n = levels[0]["nx"] * levels[0]["ny"]
    f = rng.standard_normal(n)
    fd = torch.from_numpy(f).to("cuda")
    z = torch.empty_like(fd)
    mg(fd, z)
    ref = _np_vcycle_neumann_y(levels, f[:, None], mg.coarse_sweeps, mg.omega)[:, 0]
    err = np.abs(z.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 1e-11, err


def _hierarchy_tensor_form(lat, c_elem, v_elem, sigma, min_coarse=3):
    """The hierarchy's arrays restated with tensor operations (element lookup of
    the quarter points, harmonic-mean edges, bilinear mass restriction)."""
    from qdsim_amd.lattice import LatticeMG
    L = lat
    f64 = dict(dtype=torch.float64, device="cuda")
    Y, X = torch.meshgrid(torch.arange(L.NY, **f64), torch.arange(L.NX, **f64), indexing="ij")
    xm, yy = (X[:, :-1] + 0.5) * L.hx, Y[:, :-1] * L.hy
    WE = 0.5 * (c_elem[L.element_at(xm, yy + 0.25 * L.hy)] + c_elem[L.element_at(xm, yy - 0.25 * L.hy)]) * (L.hy / L.hx)
    ym, xx = (Y[:-1, :] + 0.5) * L.hy, X[:-1, :] * L.hx
    WN = 0.5 * (c_elem[L.element_at(xx + 0.25 * L.hx, ym)] + c_elem[L.element_at(xx - 0.25 * L.hx, ym)]) * (L.hx / L.hy)
    xp, yp = X * L.hx, Y * L.hy
    vs = sum(v_elem[L.element_at(xp + sx * L.hx, yp + sy * L.hy)] for sx in (-0.25, 0.25) for sy in (-0.25, 0.25)) * 0.25
    Mterm = (vs - sigma) * (L.hx * L.hy)
    Mass = torch.full_like(Mterm, L.hx * L.hy)
    levels, NX, NY = [], L.NX, L.NY
    while True:
        d = WE[1:-1, :-1] + WE[1:-1, 1:] + WN[:-1, 1:-1] + WN[1:, 1:-1] + Mterm[1:-1, 1:-1]
        we = WE[1:-1, 1:].clone(); we[:, -1] = 0.0
        wn = WN[1:, 1:-1].clone(); wn[-1, :] = 0.0
        levels.append((d, we, wn, Mass[1:-1, 1:-1]))
        nxi, nyi = NX - 2, NY - 2
        if not (nxi % 2 == 1 and nyi % 2 == 1 and (nxi - 1) // 2 >= min_coarse and (nyi - 1) // 2 >= min_coarse):
            break
        Mass = LatticeMG._coarsen(WE, WN, Mass, NX, NY)[2]
        WE, WN, Mterm, NX, NY = LatticeMG._coarsen(WE, WN, Mterm, NX, NY)
    return levels


@pytest.mark.parametrize("order,cells,dtype", [(1, (64, 48), torch.float64), (3, (32, 32), torch.float32),
                                               (2, (40, 24), torch.float32)])
def test_hierarchy_kernels_match_tensor_form(order, cells, dtype):
    """The device-built hierarchy (qds_mg_lattice_*) equals the tensor restatement
    bitwise, level by level, including the FP64 diagonals that shift changes use."""
    from qdsim_amd.lattice import Lattice, LatticeMG
    from qdsim_amd.mesh import SimpleMesh
    m = SimpleMesh(cells[0] + 1, cells[1] + 1, 60e-9, 60e-9 * cells[1] / cells[0], order)
    lat = Lattice(m)
    rng = np.random.default_rng(3)
    c = torch.from_numpy(0.5 + 0.5 * rng.random(m.num_elements)).cuda()
    v = torch.from_numpy(0.2 * rng.random(m.num_elements) - 0.1).cuda()
    mg = LatticeMG(lat, c, v, sigma=-0.15, b=8, dtype=dtype)
    ref = _hierarchy_tensor_form(lat, c, v, -0.15)
    assert mg.n_levels == len(ref) >= 2
    for lv, (d, we, wn, mass) in zip(mg.levels_data, ref):
        assert torch.equal(lv["d64"], d)
        assert torch.equal(lv["mass"], mass)
        assert torch.equal(lv["d"], d.reshape(-1).to(dtype))
        assert torch.equal(lv["we"], we.reshape(-1).to(dtype))
        assert torch.equal(lv["wn"], wn.reshape(-1).to(dtype))
    # a shift change moves every diagonal by (old - new shift) x lumped mass
    mg.set_sigma(-0.2)
    for lv, (d, we, wn, mass) in zip(mg.levels_data, ref):
        assert torch.allclose(lv["d64"], d + 0.05 * mass, rtol=1e-14, atol=0)
