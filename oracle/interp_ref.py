"""Oracle: FEInterpolator point evaluation (TEST INFRASTRUCTURE ONLY).

Restates backend/src/fe_interpolator.cpp:
  interpolate ............. :54-160  (value; outside the mesh: nearest node's value)
  interpolateWithGradient . :162-360 (outside: inverse-distance gradient estimate over
                                      the nodes sharing an element with the nearest node)
  findElement ............. :363-410 (linear scan, first element whose barycentrics are
                                      all >= -1e-10)
  barycentrics ............ :433-455
  shape values/gradients .. :474-518, :542-613 (fem_ref.shape_values / shape_gradients)
Pure numpy, brute force over every element and node (small meshes only).
Pinned by: exact reproduction of polynomials of degree <= p (values and
gradients), partition of unity, and the brute-force nearest-node rule.
"""
import numpy as np

from . import fem_ref


def barycentric(px, py, v):
    """fe_interpolator.cpp:433-455: (inside, lambda[3]) for vertices v (3, 2)."""
    (x0, y0), (x1, y1), (x2, y2) = v
    area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    l0 = 0.5 * ((x1 * y2 - x2 * y1) + (y1 - y2) * px + (x2 - x1) * py) / area
    l1 = 0.5 * ((x2 * y0 - x0 * y2) + (y2 - y0) * px + (x0 - x2) * py) / area
    l2 = 0.5 * ((x0 * y1 - x1 * y0) + (y0 - y1) * px + (x1 - x0) * py) / area
    eps = 1e-10
    return (l0 >= -eps) and (l1 >= -eps) and (l2 >= -eps), np.array([l0, l1, l2])


def find_element(px, py, x, y, el):
    """fe_interpolator.cpp:363-410 (vectorised over elements, first hit)."""
    v0, v1, v2 = el[:, 0], el[:, 1], el[:, 2]
    x0, y0, x1, y1, x2, y2 = x[v0], y[v0], x[v1], y[v1], x[v2], y[v2]
    area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    l0 = 0.5 * ((x1 * y2 - x2 * y1) + (y1 - y2) * px + (x2 - x1) * py) / area
    l1 = 0.5 * ((x2 * y0 - x0 * y2) + (y2 - y0) * px + (x0 - x2) * py) / area
    l2 = 0.5 * ((x0 * y1 - x1 * y0) + (y0 - y1) * px + (x1 - x0) * py) / area
    eps = 1e-10
    hit = np.nonzero((l0 >= -eps) & (l1 >= -eps) & (l2 >= -eps))[0]
    return int(hit[0]) if len(hit) else -1


def interpolate(px, py, x, y, el, order, field, gradient=False):
    """Value (and gradient) of the FE field at one point, as the reference computes it."""
    e = find_element(px, py, x, y, el)
    if e < 0:
        d = (px - x) ** 2 + (py - y) ** 2
        nd = int(np.argmin(d))          # first index of the minimum, as the strict '<' scan
        if not gradient:
            return float(field[nd])
        conn = []
        for row in el:
            if nd in row:
                for j in row:
                    if j != nd and j not in conn:
                        conn.append(int(j))
        gx = gy = tw = 0.0
        for j in conn:
            dx, dy = x[j] - x[nd], y[j] - y[nd]
            dist = np.sqrt(dx * dx + dy * dy)
            if dist > 1e-10:
                w = 1.0 / dist
                df = field[j] - field[nd]
                gx += w * (df * dx / dist)
                gy += w * (df * dy / dist)
                tw += w
        if tw > 0:
            gx, gy = gx / tw, gy / tw
        else:
            gx = gy = 0.0
        return float(field[nd]), gx, gy
    nodes = el[e]
    v = np.stack([x[nodes[:3]], y[nodes[:3]]], 1)
    _, lam = barycentric(px, py, v)
    l0, l1, l2 = (np.array([t]) for t in lam)
    N = np.array([s[0] for s in fem_ref.shape_values(order, l0, l1, l2)])
    val = float(N @ field[nodes])
    if not gradient:
        return val
    (x0, y0), (x1, y1), (x2, y2) = v
    area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    g0 = np.array([0.5 * (y1 - y2) / area, 0.5 * (x2 - x1) / area])
    g1 = np.array([0.5 * (y2 - y0) / area, 0.5 * (x0 - x2) / area])
    g2 = np.array([0.5 * (y0 - y1) / area, 0.5 * (x1 - x0) / area])
    G = np.array([gk[0] for gk in fem_ref.shape_gradients(order, l0, l1, l2, g0[None], g1[None], g2[None])])
    g = G.T @ field[nodes]
    return val, float(g[0]), float(g[1])
