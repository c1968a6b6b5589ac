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
  mesh_ref.mesh_cpp_p2 ........... pinned: brute-force std::map restatement, counts
  mesh_ref.mesh_p3 ............... pinned: counts (Table of DOF/nnz counts), lattice
  fem_ref shape functions ........ pinned: partition of unity, Kronecker, gradients sum 0
  fem_ref quadrature ............. pinned: monomial exactness up to the rule degree
  fem_ref element matrices ....... pinned: closed forms (P1 K, P1/P2 M), K*1=0, sum M = area
  fem_ref.assemble ............... pinned: scipy COO->CSR canonical form, nnz counts
  eig_ref ........................ pinned: tridiag(-1,2,-1) spectrum, particle in a box
  poisson_ref .................... pinned: linear solution phi = V x / Lx (P1, rho = 0)
  cap_ref ........................ pinned: LAPACK zggev of the oracle pencil; W = 0 reduces to eig_ref
  fixed_ref ...................... pinned: restated matrices vs the restricted symmetric pencil
  interp_ref ..................... pinned: exact reproduction of degree-p polynomials (values and
                                   gradients), brute-force nearest node outside the mesh
"""

PINNING = __doc__
