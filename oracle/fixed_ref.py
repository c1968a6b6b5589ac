"""Restatement of FixedOpenSystemSolver's own formulas (oracle; test use only).

qdsim_cython/solvers/fixed_open_system_solver.pyx:85-119 (mesh), :124-142 (CAP
absorption), :197-271 (lumped element matrices), :273-291 (identity rows at nodes
0 and N-1), :293-319 (eigsh which='SM', tol=1e-6), :321-328 (constant lifetime).
Loop-based over elements, SI units."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HBAR = 1.054571817e-34
EV = 1.602176634e-19


def cap_absorption(x, y, Lx, Ly, cap_strength, cap_length_ratio):
    cap_length = cap_length_ratio * min(Lx, Ly)
    if x < cap_length:
        return cap_strength * (1.0 - x / cap_length) ** 2
    if x > Lx - cap_length:
        return cap_strength * (1.0 - (Lx - x) / cap_length) ** 2
    return 0.0


def matrices(nx, ny, Lx, Ly, m_star_func, potential_func, open_=False, cap_strength=0.01 * EV,
             cap_length_ratio=0.2):
    xs = np.array([i * Lx / (nx - 1) for j in range(ny) for i in range(nx)])
    ys = np.array([j * Ly / (ny - 1) for j in range(ny) for i in range(nx)])
    rows, cols, hd, md = [], [], [], []
    for j in range(ny - 1):
        for i in range(nx - 1):
            for tri in ([j * nx + i, j * nx + i + 1, (j + 1) * nx + i],
                        [j * nx + i + 1, (j + 1) * nx + i + 1, (j + 1) * nx + i]):
                x0, x1, x2 = xs[tri]
                y0, y1, y2 = ys[tri]
                area = 0.5 * abs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
                if area < 1e-20:
                    continue
                xc, yc = (x0 + x1 + x2) / 3.0, (y0 + y1 + y2) / 3.0
                m = m_star_func(xc, yc)
                V = potential_func(xc, yc)
                if open_:
                    V += cap_absorption(xc, yc, Lx, Ly, cap_strength, cap_length_ratio)
                kin = HBAR * HBAR / (2.0 * m) * area / 3.0
                pot = V * area / 3.0
                mf = area / 12.0
                for a in range(3):
                    for b in range(3):
                        rows.append(tri[a]); cols.append(tri[b])
                        hd.append(kin + pot if a == b else 0.5 * kin)
                        md.append(2.0 * mf if a == b else mf)
    n = nx * ny
    H = sp.csr_matrix((hd, (rows, cols)), shape=(n, n))
    M = sp.csr_matrix((md, (rows, cols)), shape=(n, n))
    return xs, ys, H, M


def reference_call(H, M, k):
    """The exact boundary treatment and eigsh call of the reference."""
    n = H.shape[0]
    Hl, Ml = H.tolil(), M.tolil()
    for node in (0, n - 1):
        Hl[node, :] = 0; Hl[node, node] = 1.0
        Ml[node, :] = 0; Ml[node, node] = 1.0
    k = max(1, min(k, n - 3))
    return spla.eigsh(Hl.tocsr(), k=k, M=Ml.tocsr(), which='SM', tol=1e-6, maxiter=1000)
