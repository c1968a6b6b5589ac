"""Closed-system Schroedinger eigenproblem on the GPU: assembly -> Dirichlet
restriction -> mixed-precision block LOBPCG.

Reference path: CythonSchrodingerSolver._assemble_matrices + solve
(qdsim_cython/solvers/schrodinger_solver.pyx:122-199, 417-526) and the P1-P3
element formulas of backend/src/fem.cpp:924-989.  Differences (kept on purpose):

* consistent FEM potential matrix int V phi_i phi_j at the quadrature points (the
  Cython solver adds V A/3 to the diagonal, twice on two of the nodes,
  schrodinger_solver.pyx:224-236);
* homogeneous Dirichlet by restriction to the interior DOFs (the reference
  replaces boundary rows by identity rows, which leaves a non-symmetric pencil and
  spurious eigenvalues at 1.0 J, schrodinger_solver.pyx:253-283);
* the lowest eigenvalues are returned (which='SA'); the reference asks ARPACK for
  which='SM', which coincides for positive spectra.

Units: SI at the API, nm / eV on the device.  Eigenvectors are M-normalised in SI
(psi^T M psi = 1 with M in m^2), as ARPACK returns them for the SI pencil.
"""
import time

import numpy as np
import torch

from . import kernels
from .device import sync
from .fem import FEMSpace
from .mesh import SimpleMesh
from .lattice import Lattice, LatticeMG
from .lobpcg import CSROperator, LatticeOperator, lobpcg, lobpcg_mixed, solve_mixed
from .units import EV_TO_J, HBAR, NM

KIN_SCALE = HBAR * HBAR / 2.0 / EV_TO_J / (NM * NM)   # hbar^2/2 in eV nm^2 kg


def default_block(k, n):
    """Block size of the block eigensolver for k wanted pairs of an n-dimensional
    pencil: k + 8 rounded up to a multiple of 4 (at least 8), reduced while 4b > n
    but never below k (rounded up to a multiple of 4).  Returns None when no such
    block has 3b <= n (the [X | W | P] basis would not fit): the caller then needs
    the dense eigensolve."""
    kmin = max(4, (k + 3) // 4 * 4)
    b = max(((k + 8) + 3) // 4 * 4, 8)
    while 4 * b > n and b - 4 >= kmin:
        b -= 4
    return b if 3 * b <= n else None


def breakdown_reason(ritz, converged, error=None):
    """Why a mixed solve must be redone with the safe shift, or None: a failed
    Rayleigh-Ritz (error), non-finite Ritz values, or no convergence."""
    if error is not None:
        return f"Rayleigh-Ritz failed: {error}"
    if ritz is None or not np.all(np.isfinite(ritz)):
        return "non-finite Ritz values"
    if not converged:
        return "not converged"
    return None


def _warm(start, n_free, b):
    """True when ``start`` (an earlier, converged solve's info) holds a block for
    this problem."""
    if not start or start.get("X_block") is None or not start.get("converged", True):
        return False
    X = start["X_block"]
    return X.dim() == 2 and X.shape[0] == n_free and X.shape[1] >= b


def shift_below(theta, floor, frac):
    """Multigrid shift just below the lowest eigenvalue: theta_1 - frac (theta_b -
    theta_1) from a block's Ritz values theta (upper bounds of the eigenvalues),
    never below floor (the shift below min V, always safe).  The block iteration
    with the preconditioner (H - sigma M)^-1 converges at a rate set by the
    relative gaps (lambda_b+1 - lambda_i) / (lambda_b+1 - sigma): on a quantum dot
    sigma below min V makes them ~0.015, a shift at theta_1 minus a tenth of the
    block's spread nearly doubles them."""
    th = np.sort(np.asarray(theta.detach().cpu().numpy() if torch.is_tensor(theta) else theta, dtype=np.float64))
    if th.size < 2 or not np.all(np.isfinite(th)):
        return floor
    return max(float(floor), float(th[0] - frac * (th[-1] - th[0])))


class SchrodingerProblem:
    """H = K(hbar^2/2m*) + M(V) [- i M(W)], M on a structured SimpleMesh."""

    def __init__(self, mesh, m_star_func=None, potential_func=None, sampling="quadrature"):
        t0 = time.perf_counter()
        self.mesh = mesh
        self.space = FEMSpace(mesh)
        self.lat = Lattice(mesh)
        r_ip, r_ix, r_src = kernels.pattern_restrict(self.space.indptr_host, self.space.indices_host,
                                                     self.lat.new_of_old, self.lat.n_free)
        dev = self.space.dev
        self.r_indptr = torch.from_numpy(r_ip).to(dev)
        self.r_indices = torch.from_numpy(r_ix).to(dev)
        self.r_src = torch.from_numpy(r_src).to(dev)
        # lattice (stencil) form of the free-DOF operator: value slot -> full CSR position
        self.stencil, slot_src = kernels.lattice_stencil(self.lat.p, self.lat.nx_i, self.lat.ny_i, r_ip, r_ix)
        s_src = np.where(slot_src >= 0, r_src[np.maximum(slot_src, 0)], -1).astype(np.int32)
        self.s_src = torch.from_numpy(s_src).to(dev)
        del slot_src, s_src
        self.use_lattice = True   # products in the lattice form (False: CSR)
        self.old_of_new = torch.from_numpy(self.lat.old_of_new).to(dev)
        self.n_free = self.lat.n_free
        self.sampling = sampling
        self.c_q = self.v_q = self.w_q = None
        self.H = self.M = self.W = None
        self.mg_options = {}          # extra LatticeMG arguments (cycle shape)
        self._m_func = self._v_func = None
        self._coarse = None           # coarse problem of the nested start (built on first use)
        self.nested = True
        # a mixed solve that breaks down or does not converge is solved again with
        # the shift below min V (info["breakdown"] names the reason, info["retried"]
        # says whether it was solved again)
        self.retry = True
        # mixed-mode iteration: "mixed" (X in FP64, corrections W, P in FP32, one
        # phase) or "two-phase" (all-FP32 phase to its floor, then all-FP64)
        self.scheme = "mixed"
        self.setup_time = time.perf_counter() - t0
        if m_star_func is not None:
            self.set_mass(m_star_func)
        if potential_func is not None:
            self.set_potential(potential_func)

    # -- coefficients at quadrature points (device, eV / eV nm^2) ----------------
    def _where(self):
        return "centroid" if self.sampling == "centroid" else "quadrature"

    def set_mass(self, m_star_func):
        m = self.space.sample(m_star_func, self._where())
        self.c_q = (KIN_SCALE / m).contiguous()
        self._m_func = m_star_func
        self._coarse = None

    def set_potential(self, potential_func, extra_ev=None):
        v = self.space.sample(potential_func, self._where(), 1.0 / EV_TO_J)
        if extra_ev is not None:
            v = v + extra_ev
        self.v_q = v.contiguous()
        self._v_func = potential_func if extra_ev is None else None
        self._coarse = None

    def set_potential_ev(self, v_q):
        self.v_q = v_q.contiguous()
        self._v_func = None
        self._coarse = None

    def set_cap_ev(self, w_q):
        self.w_q = None if w_q is None else w_q.contiguous()

    # -- assembly ---------------------------------------------------------------
    def assemble(self):
        out = self.space.assemble(c=self.c_q, v=self.v_q, stiffness_mass=True, mass=True,
                                  w=self.w_q, cap=self.w_q is not None)
        self.H, self.M, self.W = out
        return out

    def operator(self, dtype=torch.float64, lattice=None):
        """Free-DOF (H, M) in HBM: the lattice (stencil) form by default, CSR on request."""
        if self.use_lattice if lattice is None else lattice:
            H = kernels.csr_take(self.s_src, self.H, dtype=dtype)
            M = kernels.csr_take(self.s_src, self.M, dtype=dtype)
            return LatticeOperator(self.stencil, self.n_free, H, M, dtype)
        H = kernels.csr_take(self.r_src, self.H, dtype=dtype)
        M = kernels.csr_take(self.r_src, self.M, dtype=dtype)
        return CSROperator(self.r_indptr, self.r_indices, self.n_free, H, M, dtype)

    def preconditioner(self, b, dtype=torch.float32, sigma=None, cycle=None):
        """Lattice multigrid for (H - sigma M).  cycle: LatticeMG options; the
        default V(1,1) suits the mixed solve, whose FP64 phase starts close to the
        answer; an all-FP64 solve from a random start keeps V(2,2) with 40 coarse
        sweeps (with V(1,1) its Rayleigh-Ritz basis degenerated in every other
        iteration on the benchmark problem: 245 instead of 80 iterations)."""
        c_e = self.c_q.mean(dim=1) if self.c_q.dim() == 2 else self.c_q
        v_e = self.v_q.mean(dim=1) if self.v_q.dim() == 2 else self.v_q
        if sigma is None:
            vmin = float(v_e.min().item())
            vmax = float(v_e.max().item())
            sigma = vmin - 0.05 * (vmax - vmin) - 1e-3
        opts = dict(cycle or {})
        opts.update(self.mg_options)
        return LatticeMG(self.lat, c_e, v_e, sigma, b, dtype=dtype, **opts)

    # -- nested start -----------------------------------------------------------
    def _nested_start(self, k, b, factor=None, tol=None):
        """Initial block from the same problem on a mesh with factor-times fewer cells
        per direction: its FP32 block eigenvectors (tol relative residual), evaluated
        at the free DOFs of this mesh with the FE interpolation kernel.  Starting the
        fine solve there instead of from random vectors saves the iterations that
        resolve the smooth part of the eigenvectors (nested iteration).  Needs the
        mass and potential as functions and a cell count divisible by factor with
        at least 32 coarse cells; returns None otherwise.  The coarse problem (mesh,
        pattern, quadrature data) is built once and reused; every solve re-assembles
        and re-solves it.

        Default factor: 4, doubled while the coarse mesh keeps at least 64 cells
        across (512^2 cells: a 64^2 coarse mesh), with tol 1e-2 at factor 4 and
        5e-3 above.  On the 512^2-cell P3 dot the 64^2 coarse solve (18 iterations,
        14 ms) starts the fine iteration as well as the 128^2 one (19 iterations,
        18 ms): 143 against 151 ms per solve."""
        cx, cy = self.mesh.nx - 1, self.mesh.ny - 1
        if factor is None:
            factor = 4
            while min(cx, cy) // (2 * factor) >= 64 and cx % (2 * factor) == 0 and cy % (2 * factor) == 0:
                factor *= 2
        if tol is None:
            tol = 1e-2 if factor == 4 else 5e-3
        if (not self.nested or self._m_func is None or self._v_func is None or cx % factor or cy % factor
                or min(cx, cy) // factor < 32):
            return None
        if self._coarse is None:
            self._coarse = {}
        if factor not in self._coarse:
            from .interpolator import FEInterpolator
            cm = SimpleMesh(cx // factor + 1, cy // factor + 1, self.mesh.Lx, self.mesh.Ly,
                            self.mesh.element_order)
            cp = SchrodingerProblem(cm, self._m_func, self._v_func, sampling=self.sampling)
            cp.nested = False
            L = self.lat
            # free DOF coordinates of this mesh in lattice order (SI, the mesh's units)
            iy, ix = np.divmod(np.arange(L.n_free), L.nx_i)
            px = torch.from_numpy((ix + 1) * (self.mesh.Lx / (L.NX - 1))).to(self.space.dev)
            py = torch.from_numpy((iy + 1) * (self.mesh.Ly / (L.NY - 1))).to(self.space.dev)
            self._coarse[factor] = (cp, FEInterpolator(cm), px, py)
        cp, interp, px, py = self._coarse[factor]
        cp.assemble()
        op32 = cp.operator(torch.float32)
        pre = cp.preconditioner(b, torch.float32)
        sigma0 = pre.sigma
        # random start smoothed by 16 preconditioner applications (inverse iteration
        # with the shift below min V, no host round trip), shift moved to the Ritz
        # values after 4 iterations: 21 instead of 42 iterations on the 128^2-cell P3
        # dot from one application and the fixed shift
        r = lobpcg(op32, k, b, pre, tol=tol, maxiter=200, stall_iters=8, refresh=1, smooth=16,
                   retune=lambda th: pre.set_sigma(shift_below(th, sigma0, 0.25)), retune_iter=4)
        full = torch.zeros((cp.space.n, b), dtype=torch.float64, device=r.X.device)
        full[cp.old_of_new] = r.X.double()
        self._nested_iters = r.iterations
        self._nested_theta = r.eigenvalues
        return interp.interpolate_points(px, py, full)

    # -- solve ------------------------------------------------------------------
    def solve(self, k, b=None, mode="mixed", tol=1e-6, tol32=3e-3, maxiter=400, seed=0, verbose=False,
              nested_opts=None, b64=None, start=None, **mixed_opts):
        """Lowest k eigenpairs.  Returns (eigenvalues eV (np), full-numbering
        eigenvectors (n, k) float64 device tensor, M_nm-normalised, info).

        start: the info dict of an earlier solve of this problem with the same k
        and block (its final block ``X_block``): a warm start for a changed
        potential (the self-consistent loop), which replaces the nested or random
        start."""
        if self.H is None:
            self.assemble()
        if b is None:
            b = default_block(k, self.n_free)
            if b is None:
                raise ValueError(f"{k} states of a {self.n_free}-DOF problem: the block eigensolver needs "
                                 f"3 b <= n with b >= k; use a finer mesh or fewer states")
        t0 = time.perf_counter()
        op64 = self.operator(torch.float64)
        if mode == "mixed":
            pre32 = self.preconditioner(b, torch.float32)
            b64 = b if b64 is None else int(b64)
            pre32_64 = pre32 if b64 == b else self.preconditioner(b64, torch.float32)
            sync()
            t_prep = time.perf_counter() - t0

            def pre64(R, W, _pre=pre32_64, _buf={}):
                if _pre.takes_fp64 and R.stride(0) == R.shape[1]:
                    # the FP32 hierarchy reads R and writes W in FP64 directly
                    _pre(R, W)
                    return
                if "r" not in _buf:
                    _buf["r"] = torch.empty((R.shape[0], b64), dtype=torch.float32, device=R.device)
                    _buf["w"] = torch.empty((R.shape[0], b64), dtype=torch.float32, device=R.device)
                _buf["r"].copy_(R)
                _pre(_buf["r"], _buf["w"])
                W.copy_(_buf["w"])

            t_n = time.perf_counter()
            self._nested_iters = 0
            self._nested_theta = None
            sigma0 = pre32.sigma
            if _warm(start, self.n_free, b):
                # the start block's Rayleigh quotients under this operator bound its
                # eigenvalues from above (the previous solve's Ritz values do not:
                # the operator has changed since)
                X0 = start["X_block"][:, :b].contiguous()
                HX0, MX0 = torch.empty_like(X0), torch.empty_like(X0)
                op64.apply(X0, HX0, MX0)
                self._nested_theta = kernels.coldot(X0, HX0) / kernels.coldot(X0, MX0)
            else:
                X0 = self._nested_start(k, b, **(nested_opts or {}))
            if self._nested_theta is not None:
                # coarse (or previous) Ritz values: a conservative shift for the FP32 phase
                pre32.set_sigma(shift_below(self._nested_theta, sigma0, 0.25))
            sync()
            t_nested = time.perf_counter() - t_n

            def retune(r32):
                # fine Ritz values of the FP32 phase: the shift of the FP64 phase
                pre32_64.set_sigma(shift_below(r32.eigenvalues, sigma0, 0.1))

            scheme = mixed_opts.pop("scheme", self.scheme)

            def run(tuned):
                if scheme == "mixed":
                    # one phase: the shift moves once, from the Ritz values of the
                    # iteration itself, when the residual reaches the switch point
                    # of the two-phase scheme
                    t_m = time.perf_counter()
                    r = lobpcg_mixed(op64, k, b, pre32, X0, tol=tol, maxiter=maxiter, verbose=verbose,
                                     retune=(lambda th: pre32.set_sigma(shift_below(th, sigma0, 0.1))) if tuned
                                     else None, retune_at=tol32, seed=seed)
                    r.timings.update({"mixed_s": time.perf_counter() - t_m, "mixed_iters": r.iterations})
                    r.history = {"mixed": r.history}
                    r.col_history = {"mixed": r.col_history}
                    return r
                return solve_mixed(op64, k, b, pre32, pre64, tol=tol, tol32=tol32, maxiter=maxiter, seed=seed,
                                   verbose=verbose, X0=X0, b64=b64, between=retune if tuned else None, **mixed_opts)

            try:
                res = run(True)
                breakdown = breakdown_reason(res.eigenvalues.cpu().numpy(), res.converged)
            except np.linalg.LinAlgError as e:
                breakdown = breakdown_reason(None, False, e)
            retried = breakdown is not None and self.retry
            if retried:
                # a shift above the lowest eigenvalue (a block that missed the lowest
                # state, whose Ritz values then bound the wrong eigenvalues) makes the
                # multigrid operator indefinite; solve again with the shift below min V
                pre32.set_sigma(sigma0)
                pre32_64.set_sigma(sigma0)
                res = run(False)
            elif breakdown is not None and breakdown != "not converged":
                raise np.linalg.LinAlgError(f"mixed-precision solve broke down ({breakdown}) with the retry off")
        elif mode == "fp64":
            # the same start as the mixed solve (warm block or nested coarse solve), so
            # that the two modes differ in precision only
            breakdown, retried = None, False
            pre = self.preconditioner(b, torch.float64, cycle=dict(nu_pre=2, nu_post=2, coarse_sweeps=40))
            sync()
            t_prep = time.perf_counter() - t0
            t_n = time.perf_counter()
            self._nested_iters = 0
            self._nested_theta = None
            sigma0 = pre.sigma
            if _warm(start, self.n_free, b):
                X0 = start["X_block"][:, :b].contiguous()
            else:
                X0 = self._nested_start(k, b, **(nested_opts or {}))
            if self._nested_theta is not None:
                pre.set_sigma(shift_below(self._nested_theta, sigma0, 0.25))
            sync()
            t_nested = time.perf_counter() - t_n
            res = lobpcg(op64, k, b, pre, X0=X0, tol=tol, maxiter=maxiter, seed=seed, verbose=verbose)
        else:
            raise ValueError(f"unknown mode {mode}")
        lam = res.eigenvalues[:k].double()
        X = res.X[:, :k].double()
        full = torch.zeros((self.space.n, k), dtype=torch.float64, device=X.device)
        full[self.old_of_new] = X
        sync()
        info = {"iterations": res.iterations, "converged": bool(res.converged),
                "breakdown": breakdown, "retried": retried,
                "residuals": res.residuals[:k].cpu().numpy() if res.residuals is not None else None,
                "timings": dict(res.timings, prep_s=t_prep, nested_s=t_nested,
                                nested_iters=self._nested_iters),
                "history": res.history, "col_history": getattr(res, "col_history", None), "block": b, "X_block": res.X, "ritz": res.eigenvalues,
                "solve_s": time.perf_counter() - t0}
        return lam.cpu().numpy(), full, info
