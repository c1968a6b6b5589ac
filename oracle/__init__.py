"""CPU FP64 oracle for the FEM-Hamiltonian path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithms in plain numpy / scipy so the HIP
path can be checked against them.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline -- never as part of the product path.  The
product package ``qdsim_amd`` does not import anything from here.

Each function cites the reference file:line it follows.  How each one is pinned
(golden values, closed forms, invariants) is listed in ``PINNING`` below and in
the tests that exercise it.

Parity status per function:
  mesh_ref.simple_mesh_p1 ........ pinned: mesh_minimal.pyx:236-261 asserts + counts
  mesh_ref.mesh_cpp_p2 ........... pinned: hand-derived ids and coordinates of 1x1 and 2x1-cell
                                   meshes with and without the m03 orphan (mesh.cpp:95-115
                                   call order), counts
  mesh_ref.mesh_p3 ............... pinned: hand-derived ids of 1x1 and 2x1-cell meshes (a shared
                                   edge met in both orientations), counts, lattice
  fem_ref shape functions ........ pinned: partition of unity, Kronecker, gradients sum 0
  fem_ref quadrature ............. pinned: monomial exactness up to the rule degree
  fem_ref element matrices ....... pinned: closed forms (P1 K, P1/P2 M), K*1=0, sum M = area
  fem_ref.assemble ............... pinned: scipy COO->CSR canonical form, nnz counts
  eig_ref ........................ pinned: tridiag(-1,2,-1) spectrum, particle in a box
  poisson_ref .................... pinned: linear solution phi = V x / Lx (P1, rho = 0); the
                                   nodally exact quadratic for uniform rho in both sign
                                   conventions (fixes the sign and the mass weighting)
  cap_ref ........................ pinned: LAPACK zggev of the oracle pencil; W = 0 reduces to eig_ref
  fixed_ref ...................... pinned: hand-computed H and M of 2x2 and 3x2-node meshes from the
                                   lumped formulas of fixed_open_system_solver.pyx:230-247
  sc_ref ......................... pinned by composition: eig_ref and poisson_ref (each pinned
                                   above) joined by the density formulas of
                                   self_consistent_solver.pyx:257-286
  sc_adjoint_ref ................. pinned: the map at the ramp equals sc_ref.one_iteration; the
                                   dense Jacobians against central differences of the map; the
                                   reverse-mode gradients (v0, V_p, V_n, N_A, N_D) against central
                                   differences of the Newton-converged fixed point
  interp_ref ..................... pinned: exact reproduction of degree-p polynomials (values and
                                   gradients), brute-force nearest node outside the mesh
  dd_ref ......................... pinned: Poisson step = nodally exact parabola for uniform charge
                                   (sign, 1/6-weight factor, eps0, Dirichlet rows); field and currents
                                   of linear fields; Boltzmann round trip of the mass-action carriers;
                                   continuity operator: K 1 = 0, divergence identity for the drift
                                   term, E = 0 against fem_ref's stiffness, load sums; continuity
                                   step = nodally exact parabola; ldlt_lower against np.tril.  The
                                   lower-triangle reading of SimplicialLDLT is Eigen's documented
                                   default, not executed here
"""

PINNING = __doc__
