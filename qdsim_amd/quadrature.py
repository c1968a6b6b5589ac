"""Quadrature rules and Lagrange shape-function tables on the reference triangle.

The reference's P2/P3 rules are truncated to four digits or use equal weights
(backend/src/fem.cpp:861-913), so they are replaced by fully symmetric rules of
degree 2p+2 whose points and weights were solved to machine precision from the
moment equations (degree 4: 6 points, degree 6: 12 points, degree 8: 16 points;
the Dunavant families).  Exactness on every monomial up to the rule degree is a
unit test.

Shape functions are built generically: the nodal Lagrange basis on the element's
DOF positions is the inverse of the monomial Vandermonde matrix, so P1/P2/P3 share
one code path.  DOF positions follow the element DOF order of the mesh module
(fe_interpolator.cpp:474-518 ordering).
"""
import numpy as np

# orbit tables: ('c', w) | ('a', w, a) -> (a, a, 1-2a) | ('b', w, a, b) -> 6 perms of (a, b, 1-a-b)
_RULES = {
    1: [('c', 1.0)],
    2: [('a', 1.0 / 3.0, 1.0 / 6.0)],
    4: [('a', 0.22338158967801136, 0.44594849091596483),
        ('a', 0.10995174365532195, 0.09157621350977076)],
    6: [('a', 0.11678627572637744, 0.24928674517091226),
        ('a', 0.05084490637020631, 0.06308901449150205),
        ('b', 0.08285107561837479, 0.05314504984481764, 0.3103524510337826)],
    8: [('c', 0.144315607677787),
        ('a', 0.095091634267285, 0.459292588292723),
        ('a', 0.103217370534718, 0.17056930775176),
        ('a', 0.032458497623198, 0.050547228317031),
        ('b', 0.027230314174435, 0.008394777409958, 0.263112829634638)],
}


def triangle_rule(degree):
    """Barycentric points (Q, 3) and weights (Q,) summing to 1 (multiply by area)."""
    deg = min(d for d in _RULES if d >= degree)
    pts, ws = [], []
    for orb in _RULES[deg]:
        if orb[0] == 'c':
            pts.append((1 / 3, 1 / 3, 1 / 3)); ws.append(orb[1])
        elif orb[0] == 'a':
            w, a = orb[1], orb[2]
            for p in [(a, a, 1 - 2 * a), (a, 1 - 2 * a, a), (1 - 2 * a, a, a)]:
                pts.append(p); ws.append(w)
        else:
            w, a, b = orb[1], orb[2], orb[3]
            c = 1 - a - b
            for p in [(a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)]:
                pts.append(p); ws.append(w)
    return np.array(pts), np.array(ws)


def default_degree(order):
    return 2 * order + 2


def dof_positions(order):
    """Reference-triangle (xi, eta) position of each local DOF; v0=(0,0), v1=(1,0), v2=(0,1)."""
    if order == 1:
        return np.array([[0, 0], [1, 0], [0, 1]], dtype=float)
    if order == 2:
        return np.array([[0, 0], [1, 0], [0, 1], [.5, 0], [.5, .5], [0, .5]])
    t = 1.0 / 3.0
    return np.array([[0, 0], [1, 0], [0, 1],
                     [t, 0], [2 * t, 0],          # e01: a near v0, b near v1
                     [2 * t, t], [t, 2 * t],      # e12: a near v1, b near v2
                     [0, 2 * t], [0, t],          # e20: a near v2, b near v0
                     [t, t]])


def _monomials(order):
    return [(a, b) for s in range(order + 1) for a in range(s + 1) for b in [s - a]]


def shape_tables(order, xi, eta):
    """phi, dphi/dxi, dphi/deta at points (xi, eta): arrays (Q, npe)."""
    mons = _monomials(order)
    pos = dof_positions(order)
    V = np.array([[px ** a * py ** b for (a, b) in mons] for px, py in pos])   # npe x nmon
    C = np.linalg.inv(V)                                                      # mon x npe
    xi = np.asarray(xi, float)
    eta = np.asarray(eta, float)

    def pw(x, k):
        return x ** k if k > 0 else np.ones_like(x)

    def dpw(x, k):
        return k * x ** (k - 1) if k > 0 else np.zeros_like(x)

    m = np.stack([pw(xi, a) * pw(eta, b) for a, b in mons], axis=-1)
    mx = np.stack([dpw(xi, a) * pw(eta, b) for a, b in mons], axis=-1)
    my = np.stack([pw(xi, a) * dpw(eta, b) for a, b in mons], axis=-1)
    return m @ C, mx @ C, my @ C


def element_tables(order, degree=None):
    """Tables consumed by the assembly kernel.

    Returns dict with bary (Q,3), w (Q,), phi (Q,npe), dxi (Q,npe), deta (Q,npe).
    """
    if degree is None:
        degree = default_degree(order)
    bary, w = triangle_rule(degree)
    xi, eta = bary[:, 1], bary[:, 2]
    phi, dxi, deta = shape_tables(order, xi, eta)
    return {'bary': bary, 'w': w, 'phi': phi, 'dxi': dxi, 'deta': deta, 'degree': degree}
