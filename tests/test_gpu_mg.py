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


def _mg(n_cells=64, b=8, dtype=torch.float64):
    from qdsim_amd.lattice import Lattice, LatticeMG
    from qdsim_amd.mesh import SimpleMesh
    m = SimpleMesh(n_cells + 1, n_cells + 1, 50e-9, 50e-9, 1)
    lat = Lattice(m)
    E = m.num_elements
    dev = torch.device("cuda")
    rng = np.random.default_rng(0)
    c = torch.from_numpy(0.5 + 0.5 * rng.random(E)).to(dev)
    v = torch.from_numpy(0.1 * rng.random(E)).to(dev)
    return lat, LatticeMG(lat, c, v, sigma=-0.01, b=b, dtype=dtype)


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


def test_vcycle_converges():
    from qdsim_amd import _lib, kernels
    lat, mg = _mg(128)
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
    assert factor < 0.35
