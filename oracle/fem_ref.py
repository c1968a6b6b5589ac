"""FP64 restatement of the reference element matrices and CSR assembly (oracle).

Shape functions: backend/src/fe_interpolator.cpp:433-455 (barycentrics),
:474-518 (values P1-P3), :542-613 (gradients via grad(lambda)).
Element matrices: backend/src/fem.cpp:924-989 -- per quadrature point,
H_e += w (hbar^2/2m(x_q) grad(phi_i).grad(phi_j) + V(x_q) phi_i phi_j),
M_e += w phi_i phi_j.  The quadrature rules are symmetric rules of degree 2p+2
(the reference's own tables are truncated, fem.cpp:861-913); the orbit table is
repeated here and pinned by the monomial-exactness test.
P1 centroid formulas of the Cython surface: fem_solver.pyx:277-320
(K = c A grad(lambda_i).grad(lambda_j), M = rho A/12 (1 + delta_ij)).
Global assembly: scipy COO -> CSR (fem_solver.pyx:142-145), which sums duplicate
triplets and sorts column indices.
"""
import numpy as np
import scipy.sparse as sp

_ORBITS = {
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


def rule(degree):
    """Symmetric triangle rule: barycentric points (Q,3), weights summing to 1."""
    d = min(k for k in _ORBITS if k >= degree)
    P, W = [], []
    for o in _ORBITS[d]:
        if o[0] == 'c':
            P.append((1 / 3, 1 / 3, 1 / 3)); W.append(o[1])
        elif o[0] == 'a':
            a = o[2]
            P += [(a, a, 1 - 2 * a), (a, 1 - 2 * a, a), (1 - 2 * a, a, a)]; W += [o[1]] * 3
        else:
            a, b = o[2], o[3]; c = 1 - a - b
            P += [(a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)]; W += [o[1]] * 6
    return np.array(P), np.array(W)


def shape_values(order, l0, l1, l2):
    """fe_interpolator.cpp:474-518.  Returns list of npe arrays."""
    if order == 1:
        return [l0, l1, l2]
    if order == 2:
        return [l0 * (2 * l0 - 1), l1 * (2 * l1 - 1), l2 * (2 * l2 - 1),
                4 * l0 * l1, 4 * l1 * l2, 4 * l2 * l0]
    return [0.5 * l0 * (3 * l0 - 1) * (3 * l0 - 2),
            0.5 * l1 * (3 * l1 - 1) * (3 * l1 - 2),
            0.5 * l2 * (3 * l2 - 1) * (3 * l2 - 2),
            4.5 * l0 * l1 * (3 * l0 - 1), 4.5 * l0 * l1 * (3 * l1 - 1),
            4.5 * l1 * l2 * (3 * l1 - 1), 4.5 * l1 * l2 * (3 * l2 - 1),
            4.5 * l2 * l0 * (3 * l2 - 1), 4.5 * l2 * l0 * (3 * l0 - 1),
            27 * l0 * l1 * l2]


def shape_gradients(order, l0, l1, l2, g0, g1, g2):
    """fe_interpolator.cpp:542-613; g_k = grad(lambda_k) as (..., 2) arrays."""
    l0 = l0[..., None]; l1 = l1[..., None]; l2 = l2[..., None]
    if order == 1:
        return [g0 + 0 * l0, g1 + 0 * l1, g2 + 0 * l2]
    if order == 2:
        return [g0 * (4 * l0 - 1), g1 * (4 * l1 - 1), g2 * (4 * l2 - 1),
                4 * (g0 * l1 + g1 * l0), 4 * (g1 * l2 + g2 * l1), 4 * (g2 * l0 + g0 * l2)]
    return [g0 * (0.5 * (27 * l0 * l0 - 18 * l0 + 2)),
            g1 * (0.5 * (27 * l1 * l1 - 18 * l1 + 2)),
            g2 * (0.5 * (27 * l2 * l2 - 18 * l2 + 2)),
            4.5 * (g0 * l1 * (3 * l0 - 1) + l0 * l1 * 3 * g0 + g1 * l0 * (3 * l0 - 1)),
            4.5 * (g0 * l1 * (3 * l1 - 1) + g1 * l0 * (3 * l1 - 1) + l0 * l1 * 3 * g1),
            4.5 * (g1 * l2 * (3 * l1 - 1) + l1 * l2 * 3 * g1 + g2 * l1 * (3 * l1 - 1)),
            4.5 * (g1 * l2 * (3 * l2 - 1) + g2 * l1 * (3 * l2 - 1) + l1 * l2 * 3 * g2),
            4.5 * (g2 * l0 * (3 * l2 - 1) + l2 * l0 * 3 * g2 + g0 * l2 * (3 * l2 - 1)),
            4.5 * (g2 * l0 * (3 * l0 - 1) + g0 * l2 * (3 * l0 - 1) + l2 * l0 * 3 * g0),
            27 * (g0 * l1 * l2 + l0 * g1 * l2 + l0 * l1 * g2)]


def element_geometry(vx, vy):
    """Signed area and grad(lambda_k) per element (fe_interpolator.cpp:588-600)."""
    x0, x1, x2 = vx[:, 0], vx[:, 1], vx[:, 2]
    y0, y1, y2 = vy[:, 0], vy[:, 1], vy[:, 2]
    area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    g0 = np.stack([0.5 * (y1 - y2) / area, 0.5 * (x2 - x1) / area], -1)
    g1 = np.stack([0.5 * (y2 - y0) / area, 0.5 * (x0 - x2) / area], -1)
    g2 = np.stack([0.5 * (y0 - y1) / area, 0.5 * (x1 - x0) / area], -1)
    return area, g0, g1, g2


def quad_points(vx, vy, degree):
    """Physical quadrature points (E, Q) for a rule of given degree."""
    P, W = rule(degree)
    xq = vx @ P.T
    yq = vy @ P.T
    return xq, yq


def element_matrices(order, vx, vy, degree, c_q=None, v_q=None, rho_q=None):
    """Per-element K (with coefficient c_q), V-weighted mass (v_q) and mass (rho_q).

    vx, vy: (E, 3) vertex coordinates.  c_q, v_q, rho_q: (E, Q) values at the
    quadrature points of ``rule(degree)`` or None (term skipped / rho = 1).
    Returns K, Vm, Mm with shape (E, npe, npe).
    """
    P, W = rule(degree)
    area, g0, g1, g2 = element_geometry(vx, vy)
    A = np.abs(area)
    E = vx.shape[0]
    npe = {1: 3, 2: 6, 3: 10}[order]
    K = np.zeros((E, npe, npe)); Vm = np.zeros((E, npe, npe)); Mm = np.zeros((E, npe, npe))
    for q in range(len(W)):
        l0 = np.full(E, P[q, 0]); l1 = np.full(E, P[q, 1]); l2 = np.full(E, P[q, 2])
        phi = np.stack(shape_values(order, l0, l1, l2), -1)           # E x npe
        grad = np.stack(shape_gradients(order, l0, l1, l2, g0, g1, g2), 1)  # E x npe x 2
        wq = W[q] * A
        gg = np.einsum('eik,ejk->eij', grad, grad)
        pp = np.einsum('ei,ej->eij', phi, phi)
        cq = 1.0 if c_q is None else c_q[:, q]
        K += (wq * cq)[:, None, None] * gg
        if v_q is not None:
            Vm += (wq * v_q[:, q])[:, None, None] * pp
        rq = 1.0 if rho_q is None else rho_q[:, q]
        Mm += (wq * rq)[:, None, None] * pp
    return K, Vm, Mm


def p1_centroid_matrices(vx, vy, coef, rho):
    """fem_solver.pyx:277-320 -- coef, rho: (E,) values at the element centroid."""
    area, g0, g1, g2 = element_geometry(vx, vy)
    A = np.abs(area)
    G = np.stack([g0, g1, g2], 1)
    K = (coef * A)[:, None, None] * np.einsum('eik,ejk->eij', G, G)
    M = (rho * A / 12.0)[:, None, None] * (np.ones((3, 3)) + np.eye(3))[None]
    return K, M


def assemble(conn, n, Ke):
    """COO triplets -> canonical CSR (fem_solver.pyx:142-145)."""
    E, npe = conn.shape
    rows = np.repeat(conn, npe, axis=1).ravel()
    cols = np.tile(conn, (1, npe)).ravel()
    A = sp.csr_matrix((Ke.reshape(-1), (rows, cols)), shape=(n, n))
    return A


def dirichlet_restrict(A, free):
    """Submatrix on free DOFs (the build's homogeneous-Dirichlet treatment)."""
    return A[free][:, free].tocsr()
