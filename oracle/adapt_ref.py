"""CPU restatement of the adaptive-refinement indicator.  TEST INFRASTRUCTURE ONLY:
imported by tests/ as the checker, never by the product path.

gradient_norms follows AdaptiveMesh::computeRefinementFlags
(backend/src/adaptive_mesh.cpp:431-475): per element, the pseudo-inverse of
A = [1 x y] over the element's nodes (npe rows) applied to the nodal values,
rows 1-2 of the result being the fitted gradient; the flag is |gradient| >
threshold.  Pinned (tests/test_adapt.py) by the closed form for a linear field
(exact gradient on every element, any order) and by agreement with numpy's
least-squares solver.
"""
import numpy as np


def gradient_norms(x, y, conn, psi):
    """|fitted gradient| per element; x, y, psi nodal arrays, conn (E, npe)."""
    out = np.empty(len(conn))
    for e, ce in enumerate(conn):
        A = np.column_stack([np.ones(len(ce)), x[ce], y[ce]])
        coef = np.linalg.pinv(A) @ psi[ce]
        out[e] = np.hypot(coef[1], coef[2])
    return out


def refinement_flags(x, y, conn, psi, threshold):
    return gradient_norms(x, y, conn, psi) > threshold
