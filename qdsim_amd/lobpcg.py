"""Block LOBPCG for the lowest eigenpairs of H x = lambda M x on the GPU.

Reference: the reference solves this problem with ARPACK (scipy eigsh, which='SM',
qdsim_cython/solvers/eigen_solver.pyx:225-235; shift-invert for large N,
qdsim_cython/advanced_eigenvalue_solvers.py:106-122) and with an unseeded scipy
lobpcg (eigen_solver.pyx:254-266).  This is a fresh block solver built on the
qdsim_hip kernels:

* the operator apply is the fused CSR SpMM (H and M share the pattern, one pass);
* the block iterate is S = [X | W | P] kept as three contiguous row-major (n, b)
  buffers, with H S and M S alongside; the Rayleigh-Ritz Gram matrices
  S^T [HS | MS] come out of one MFMA FP64 kernel call over the column-block lists
  (upper tiles only);
* Rayleigh-Ritz uses SVQB on the small FP64 Gram pair: scale, drop directions whose
  M-Gram eigenvalue falls below a threshold, solve the standard problem, which is
  robust when W and P become nearly dependent close to convergence;
* X, P and their H/M images are updated by the MFMA block-update kernel, writing
  into the second half of a double buffer.

Mixed precision: blocks and matrices in FP32 for the bulk of the iterations, Gram
matrices and the small eigenproblem always in FP64, then a refinement in FP64
storage started from the FP32 subspace (``solve_mixed``).
"""
import math
import time

import numpy as np
import scipy.linalg as sla
import scipy.linalg.lapack as sla_lapack
import torch

from . import kernels
from .device import blas_single_thread, device, sync


class CSROperator:
    """H and M on one CSR pattern in HBM, values in one storage dtype."""

    def __init__(self, indptr, indices, n, H, M, dtype):
        self.indptr = indptr
        self.indices = indices
        self.n = int(n)
        self.nnz = int(indices.numel())
        self.H = H.to(dtype).contiguous()
        self.M = M.to(dtype).contiguous()
        self.dtype = dtype

    def apply(self, X, HX, MX):
        kernels.spmm(self.indptr, self.indices, self.n, self.nnz, self.H, X, HX, self.M, MX)

    def apply_residual(self, X, HX, MX, lam, R, colscale, out):
        """apply, then the residual step (residual_norms) from the stored products."""
        self.apply(X, HX, MX)
        return kernels.residual_norms(HX, MX, lam, R, colscale, out=out)

    def apply_h(self, X, HX):
        kernels.spmm(self.indptr, self.indices, self.n, self.nnz, self.H, X, HX)

    def apply_m(self, X, MX):
        kernels.spmm(self.indptr, self.indices, self.n, self.nnz, self.M, X, MX)

    def astype(self, dtype):
        return CSROperator(self.indptr, self.indices, self.n, self.H, self.M, dtype)


class LatticeOperator:
    """H and M of a structured-mesh problem in the lattice (stencil) form: per-class
    value arrays without column indices, products by qds_lattice_spmm_tuned with the
    autotuned block shape (bitwise equal to the CSR product of the same matrices
    for every shape)."""

    def __init__(self, stencil, n, H, M, dtype):
        self.stencil = stencil
        self.n = int(n)
        self.nnz = int(stencil.n_slots)
        self.H = H.to(dtype).contiguous()
        self.M = M.to(dtype).contiguous()
        self.dtype = dtype

    def apply(self, X, HX, MX):
        kernels.lattice_spmm_auto(self.stencil, self.H, X, HX, self.M, MX)

    def apply_residual(self, X, HX, MX, lam, R, colscale, out):
        """apply with the residual step of the block iteration in the product's
        epilogue (qds_lattice_spmm_residual): R = (HX - MX diag(lam)) diag(colscale)
        and the column norms (views of ``out``) of residual_norms, without reading
        HX and MX back.  The block shape is the one the autotuner chose for apply."""
        from . import autotune
        ch = autotune.choice_for(self.stencil, X.dtype, X.shape[1], True, 1,
                                 lambda c: kernels.lattice_spmm_tuned(self.stencil, self.H, X, HX, self.M, MX, c))
        return kernels.lattice_spmm_residual(self.stencil, self.H, self.M, X, HX, MX, lam, R, colscale, out, ch)

    def apply_h(self, X, HX):
        kernels.lattice_spmm_auto(self.stencil, self.H, X, HX)

    def apply_m(self, X, MX):
        kernels.lattice_spmm_auto(self.stencil, self.M, X, MX)

    def astype(self, dtype):
        return LatticeOperator(self.stencil, self.n, self.H, self.M, dtype)


def _svqb_rr(G, m, nb, drop_tol):
    """Rayleigh-Ritz on S with Gram G = S^T [HS | MS] (FP64, from the MFMA kernel).

    The (m x 2m) Gram pair is at most 96 x 192 doubles; its two symmetric
    eigenproblems run in host LAPACK (dsyevd), which measured 3.5x faster than
    the device eigh for these sizes (0.72 vs 2.5 ms including the transfers) and is
    robust for the nearly rank-deficient pairs met close to convergence.
    Returns (theta[:nb], C (m x nb) as host FP64 arrays, number of dropped
    non-zero directions)."""
    g = G if isinstance(G, np.ndarray) else G.detach().cpu().numpy()
    GH = 0.5 * (g[:, :m] + g[:, :m].T)
    GM = 0.5 * (g[:, m:2 * m] + g[:, m:2 * m].T)
    dg = np.diag(GM).copy()
    # columns that vanished (locked residuals) get weight 0 and are dropped below
    ok = dg > 1e-30 * dg.max()
    n_zero = int((~ok).sum())
    d = np.where(ok, 1.0 / np.sqrt(np.where(ok, dg, 1.0)), 0.0)
    GMs = GM * d[:, None] * d[None, :]
    GHs = GH * d[:, None] * d[None, :]
    mu, Q = np.linalg.eigh(GMs)
    # keep well-conditioned directions, but never fewer than nb (eigh sorts ascending)
    n_ok = int((mu > drop_tol * mu[-1]).sum())
    nk = max(n_ok, nb)
    mu_k = np.maximum(mu[-nk:], 1e-300)
    Qk = Q[:, -nk:] / np.sqrt(mu_k)[None, :]
    Hk = Qk.T @ GHs @ Qk
    Hk = 0.5 * (Hk + Hk.T)
    theta, Y = np.linalg.eigh(Hk)
    C = d[:, None] * (Qk @ Y[:, :nb])
    # directions dropped beyond the zero columns: the basis was near-dependent
    dropped = (m - n_zero) - n_ok
    return theta[:nb], C, dropped


def _chol_rr(G, m, nb, drop_tol):
    """Rayleigh-Ritz through a Cholesky factor of the scaled M-Gram when it is well
    conditioned (reciprocal condition estimate above drop_tol): one Cholesky, three
    triangular solves and one symmetric eigensolve, about half the host cost of the
    two eigensolves of SVQB.  Returns None when the basis
    is too close to dependent; the caller then uses SVQB.  Columns that vanished
    (locked residuals) are left out and get zero weight.  This runs on the host in
    every iteration while the GPU waits, so it calls LAPACK directly and avoids
    fancy indexing when no column vanished."""
    g = G if isinstance(G, np.ndarray) else G.detach().cpu().numpy()
    GH = g[:, :m]
    GM = g[:, m:2 * m]
    dg = np.diag(GM).copy()
    keep = dg > 1e-30 * dg.max()
    if keep.all():
        idx = None
        d = 1.0 / np.sqrt(dg)
    else:
        idx = np.nonzero(keep)[0]
        if len(idx) < nb:
            return None
        d = 1.0 / np.sqrt(dg[idx])
        GH = GH.take(idx, 0).take(idx, 1)
        GM = GM.take(idx, 0).take(idx, 1)
    dd = d[:, None] * d[None, :]
    GMs = (GM + GM.T) * (0.5 * dd)
    GHs = (GH + GH.T) * (0.5 * dd)
    Lc, info = sla_lapack.dpotrf(GMs, lower=1, clean=1)
    if info != 0:
        return None
    rcond, info = sla_lapack.dpocon(Lc, float(np.abs(GMs).sum(0).max()), uplo="L")
    if info != 0 or rcond < drop_tol:
        return None
    # triangular solves, not an explicit inverse: the factor may be conditioned
    # down to drop_tol, where the inverse would cost digits of the Ritz vectors
    T, info = sla_lapack.dtrtrs(Lc, GHs, lower=1)
    T, info2 = sla_lapack.dtrtrs(Lc, np.ascontiguousarray(T.T), lower=1)
    if info != 0 or info2 != 0:
        return None
    T = 0.5 * (T + T.T)
    theta, Y, info = sla_lapack.dsyevd(T, compute_v=1, lower=1)
    if info != 0:
        return None
    theta, Y = theta[:nb], Y[:, :nb]
    Cs, info = sla_lapack.dtrtrs(Lc, Y, lower=1, trans=1)
    if info != 0:
        return None
    Cs = Cs * d[:, None]
    if idx is None:
        return theta, Cs, 0
    C = np.zeros((m, nb))
    C[idx] = Cs
    return theta, C, 0


def _rayleigh_ritz(G, m, nb, drop_tol):
    """Rayleigh-Ritz of one iteration: the library's host solver
    (qds_host_rayleigh_ritz: Cholesky of the scaled M-Gram, tridiagonalisation,
    bisection, inverse iteration; ~3-5x faster than the LAPACK calls of _chol_rr
    for m = 72) when the basis is well conditioned, SVQB otherwise."""
    out = kernels.host_rayleigh_ritz(G, m, nb, drop_tol)
    if out is not None:
        return out[0], out[1], 0
    return _svqb_rr(G, m, nb, drop_tol)


class _Staging:
    """Page-locked host buffer and its device twin for the small per-iteration
    transfers (Rayleigh-Ritz coefficients and eigenvalues, locking mask): one
    asynchronous copy each, ordered on the stream before the kernels that read it."""

    def __init__(self, size, dev):
        self.host = torch.empty(size, dtype=torch.float64, pin_memory=dev.type == "cuda")
        self.dev = torch.empty(size, dtype=torch.float64, device=dev)

    def put(self, arrays):
        """Copy host arrays back to back; returns the device views in order."""
        h = self.host.numpy()
        off, views = 0, []
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64)
            h[off:off + a.size] = a.ravel()
            views.append((off, a.shape))
            off += a.size
        self.dev[:off].copy_(self.host[:off], non_blocking=True)
        return [self.dev[o:o + int(np.prod(sh))].view(*sh) for o, sh in views]


class LOBPCGResult:
    def __init__(self, lam, X, res, iters, history, converged, timings):
        self.eigenvalues = lam
        self.X = X
        self.residuals = res
        self.iterations = iters
        self.history = history
        self.converged = converged
        self.timings = timings


def lobpcg(op, k, b, precond=None, X0=None, tol=1e-6, maxiter=500, drop_tol=None, seed=0,
           verbose=False, stall_iters=None, refresh=None, rr="host", smooth=1, retune=None, retune_iter=None):
    """rr: "host" (Rayleigh-Ritz on the host after one device-to-host copy per
    iteration) or "device" (qds_rayleigh_ritz on the GPU; the host only reads each
    iteration's status one iteration behind).  smooth: preconditioner applications
    to the seeded random start (no X0): each one is a step of inverse iteration
    with the shifted operator, which concentrates the start on the low spectrum.
    retune(theta): optional callable run once after iteration retune_iter with the
    Ritz values (host path), e.g. to move the preconditioner's shift once the Ritz
    values bound the wanted eigenvalues.  For one problem the host path is
    faster: the single-workgroup kernel of the device path takes ~0.58 ms for the
    72 x 72 projected pencil (a chain of ~5m dependent steps) against ~0.3 ms of
    single-threaded LAPACK plus a ~25 us round trip; the device path pays off for
    batches (one launch for all problems)."""
    if rr == "device" and 3 * b <= 96:
        return _lobpcg_device(op, k, b, precond, X0, tol, maxiter, drop_tol, seed, verbose, stall_iters, refresh)
    # the host Rayleigh-Ritz works on matrices of at most 96 x 192: multithreaded
    # BLAS only adds thread wake-ups there (measured 2-70x slower triangular solves)
    with blas_single_thread():
        return _lobpcg(op, k, b, precond, X0, tol, maxiter, drop_tol, seed, verbose, stall_iters, refresh, smooth,
                       retune, retune_iter)


def _default_drop_tol(dt):
    # directions whose M-Gram eigenvalue is below drop_tol carry rounding of the
    # stored vectors amplified by drop_tol^-1/2 into the projected matrices.
    # FP32: 1e-7 keeps that amplification near 2e-4 while dropping (and the P
    # restart it forces) stays rare; looser thresholds (1e-5, 1e-6) dropped a
    # direction every other iteration on the quantum-dot problems and slowed
    # convergence 1.3-2x.
    return 1e-7 if dt == torch.float32 else 1e-10


class _StallTest:
    """Stagnation close to the target (the first iterations from a random start are
    not monotone); far from it only a long stall stops."""

    def __init__(self, tol, stall_iters):
        self.tol, self.n = tol, stall_iters
        self.best = self.min_seen = math.inf
        self.best_it = self.min_it = 0

    def stalled(self, it, worst):
        if self.n is None:
            return False
        if worst < self.min_seen:
            self.min_seen, self.min_it = worst, it
        if worst < 100.0 * self.tol:
            if worst < 0.7 * self.best:
                self.best, self.best_it = worst, it
            elif it - self.best_it >= self.n:
                return True
        elif it - self.min_it >= 4 * self.n:
            return True
        return False


def _lobpcg_device(op, k, b, precond, X0, tol, maxiter, drop_tol, seed, verbose, stall_iters, refresh):
    """The block iteration of ``_lobpcg`` with the Rayleigh-Ritz step on the device
    (qds_rayleigh_ritz): residual norms -> preconditioner -> product -> Gram ->
    Rayleigh-Ritz (which also applies the residual test, the locking mask for the
    next residual and, once converged, keeps the block) -> updates, all queued on
    the stream without a host round trip.  The basis is always [X | W | P]; an
    empty or restarted P block enters as zero columns, which the Rayleigh-Ritz
    kernel leaves out.  Every iteration ends with an asynchronous copy of its status
    to page-locked memory; the host queues iteration i + 1 before it reads the
    status of iteration i, so the GPU never waits for the host, and stops one
    iteration after the status says converged (that iteration keeps the block)."""
    dev = op.H.device
    dt = op.dtype
    n = op.n
    assert b % 4 == 0 and b >= k and 3 * b <= n, (n, b, k)
    if drop_tol is None:
        drop_tol = _default_drop_tol(dt)
    t_start = time.perf_counter()

    def blk():
        return torch.zeros((n, b), dtype=dt, device=dev)

    f64 = dict(dtype=torch.float64, device=dev)
    X, HX, MX = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    P, HP, MP = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    W, HW, MW, R = blk(), blk(), blk(), blk()
    m = 3 * b
    nst = kernels.rr_status_len(b)
    G = torch.zeros((1, m, 2 * m), **f64)
    norms = torch.zeros((1, 2 * b), **f64)
    lam = torch.zeros((1, b), **f64)
    C = torch.zeros((1, m, b), **f64)
    CP = torch.zeros((1, m, b), **f64)
    colscale = torch.ones((1, b), **f64)
    status = torch.zeros((1, nst), **f64)
    status_h = torch.zeros((2, nst), dtype=torch.float64, pin_memory=dev.type == "cuda")
    events = [torch.cuda.Event(), torch.cuda.Event()] if dev.type == "cuda" else None
    c = 0
    if X0 is None:
        g = torch.Generator(device=dev).manual_seed(seed)
        X[c].copy_(torch.rand((n, b), generator=g, dtype=dt, device=dev) - 0.5)
        if precond is not None:
            R.copy_(X[c])
            precond(R, X[c])
    else:
        X[c].copy_(X0[:, :b].to(dt))
    # initial Rayleigh-Ritz on X alone
    op.apply(X[c], HX[c], MX[c])
    G0 = torch.zeros((1, b, 2 * b), **f64)
    C0 = torch.zeros((1, b, b), **f64)
    kernels.gram_blocks([X[c]], [HX[c]], [MX[c]], sym=True, out=G0[0])
    kernels.rayleigh_ritz(G0, lam, C0, status, k, tol, drop_tol)
    for src in (X, HX, MX):
        kernels.block_update_blocks([src[c]], C0[0], src[1 - c])
    c = 1 - c
    CPb = CP[0][b:]

    def iterate(c):
        kernels.residual_norms(HX[c], MX[c], lam[0], R, colscale[0], out=norms[0])
        if precond is not None:
            precond(R, W)
        else:
            W.copy_(R)
        op.apply(W, HW, MW)
        S, HS, MS = [X[c], W, P[c]], [HX[c], HW, HP[c]], [MX[c], MW, MP[c]]
        kernels.gram_blocks(S, HS, MS, sym=True, out=G[0])
        kernels.rayleigh_ritz(G, lam, C, status, k, tol, drop_tol, norms=norms, CP=CP, colscale=colscale,
                              freeze=True)
        kernels.block_update_blocks(S, C[0], X[1 - c], CP[0], P[1 - c], c2_first=b)
        if refresh:
            # images of X by product below; those of P from the W and P blocks only
            # (the X rows of the P coefficients are zero)
            for src, pimg in ((HS, HP), (MS, MP)):
                kernels.block_update_blocks(src[1:], CPb, pimg[1 - c])
            op.apply(X[1 - c], HX[1 - c], MX[1 - c])
        else:
            for src, img, pimg in ((HS, HX, HP), (MS, MX, MP)):
                kernels.block_update_blocks(src, C[0], img[1 - c], CP[0], pimg[1 - c], c2_first=b)

    def launch(it, c):
        iterate(c)
        slot = it % 2
        status_h[slot].copy_(status[0], non_blocking=True)
        if events is not None:
            events[slot].record()

    history, restarts = [], 0
    converged = False
    stall = _StallTest(tol, stall_iters)
    res = None
    t_iter = time.perf_counter()
    launch(1, c)
    c = 1 - c
    it = 0
    for it in range(1, maxiter + 1):
        if it < maxiter:
            launch(it + 1, c)
            c = 1 - c
        if events is not None:
            events[it % 2].synchronize()
        st = status_h[it % 2].numpy().copy()
        worst = float(st[0])
        history.append(worst)
        restarts += int(st[2] > 0)
        res = st[8:8 + b]
        if verbose:
            print(f"  lobpcg[{str(dt)[6:]}/dev] it {it:4d} worst rel. residual {worst:.3e}")
        if st[1] > 0:
            converged = True
            break
        if stall.stalled(it, worst):
            break
    sync()
    t_end = time.perf_counter()
    timings = {"setup_s": t_iter - t_start, "iterate_s": t_end - t_iter, "restarts": restarts}
    return LOBPCGResult(lam[0].clone(), X[c], torch.from_numpy(res.copy()), it, history, converged, timings)


def _lobpcg(op, k, b, precond=None, X0=None, tol=1e-6, maxiter=500, drop_tol=None, seed=0,
            verbose=False, stall_iters=None, refresh=None, smooth=1, retune=None, retune_iter=None):
    """Lowest k eigenpairs of (op.H, op.M) with block size b >= k.

    precond(R, W) writes W = T R (R, W are (n, b) blocks of op.dtype).
    Stops when the relative residual ||H x - lambda M x|| / (s ||M x||),
    s = max(|lambda|, max_j<k |lambda_j|), of the first k columns is below tol, or
    after maxiter iterations, or (stall_iters) when the worst residual, once below
    100 tol, has not improved by 30% for that many iterations.  Columns whose residual is below tol are
    soft-locked (their preconditioned residual is zero, so they leave the search
    space), and the P block is dropped for one iteration whenever the Rayleigh-Ritz
    basis turned out numerically dependent.
    """
    dev = op.H.device
    dt = op.dtype
    n = op.n
    assert b % 4 == 0 and b >= k and 3 * b <= n, (n, b, k)
    if drop_tol is None:
        # directions whose M-Gram eigenvalue is below drop_tol carry rounding of the
        # stored vectors amplified by drop_tol^-1/2 into the projected matrices.
        # FP32: 1e-7 keeps that amplification near 2e-4 while dropping (and the P
        # restart it forces) stays rare; looser thresholds (1e-5, 1e-6) dropped a
        # direction every other iteration on the quantum-dot problems and slowed
        # convergence 1.3-2x.
        drop_tol = 1e-7 if dt == torch.float32 else 1e-10
    t_start = time.perf_counter()

    def blk():
        # not cleared: every block is written before it is read (P and its images
        # enter the basis only after the first update wrote them), and a fill of
        # the 16 blocks cost ~1.4 ms per FP64 solve at 2.4 M DOF
        return torch.empty((n, b), dtype=dt, device=dev)

    # every block is its own contiguous (n, b) buffer: a row of a block is b
    # consecutive elements, so the block kernels read whole cache lines
    # (X, P and their images are double-buffered: the update reads the old set)
    X, HX, MX = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    P, HP, MP = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    W, HW, MW, R = blk(), blk(), blk(), blk()
    # per iteration one copy to the device: Rayleigh-Ritz coefficients,
    # eigenvalues and the locking mask of the next residual
    stage_rr = _Staging(2 * 3 * b * b + 2 * b, dev)
    stage_cs = _Staging(b, dev)
    # residual norms and the Gram of an iteration come back in one copy: device
    # buffer [norms (2b) | Gram (m x 2m)] and its page-locked host twin
    xfer = torch.empty(2 * b + 18 * b * b, dtype=torch.float64, device=dev)
    xfer_h = torch.empty(xfer.numel(), dtype=torch.float64, pin_memory=dev.type == "cuda")

    def rr_x(c):
        # Rayleigh-Ritz on the X block alone (the host solver, SVQB when the block
        # is near dependent): X, HX, MX <- their Ritz combinations in the other
        # half of the double buffer
        op.apply(X[c], HX[c], MX[c])
        G = kernels.gram_blocks([X[c]], [HX[c]], [MX[c]], sym=True)
        theta, C, _ = _rayleigh_ritz(G.cpu().numpy(), b, b, drop_tol)
        C_d, lam = stage_rr.put([C, theta])
        for src in (X, HX, MX):
            kernels.block_update_blocks([src[c]], C_d, src[1 - c])
        return 1 - c, theta, lam

    c = 0
    if X0 is None:
        g = torch.Generator(device=dev).manual_seed(seed)
        X[c].copy_(torch.rand((n, b), generator=g, dtype=dt, device=dev) - 0.5)
        if precond is not None:
            nsm = max(int(smooth), 1)
            for s_ in range(nsm):
                R.copy_(X[c])
                precond(R, X[c])
                if (s_ + 1) % 4 == 0 and s_ + 1 < nsm:
                    # subspace iteration: four steps of inverse iteration pull the
                    # columns towards the lowest eigenvectors by up to
                    # ((lambda_1 - sigma) / (lambda_b - sigma))^4; the Rayleigh-Ritz
                    # step re-orthonormalises before the basis loses rank in FP32
                    c, _, _ = rr_x(c)
    else:
        X[c].copy_(X0[:, :b].to(dt))
    # initial Rayleigh-Ritz on X alone
    c, theta, lam = rr_x(c)
    lam_h = theta
    have_p = False
    history = []
    col_history = []   # relative residual of every block column, per iteration
    drops = []
    converged = False
    it = 0
    best = min_seen = math.inf
    best_it = min_it = 0
    res = None
    t_iter = time.perf_counter()
    (colscale,) = stage_cs.put([np.ones(b)])
    for it in range(1, maxiter + 1):
        kernels.residual_norms(HX[c], MX[c], lam, R, colscale, out=xfer[:2 * b])
        # the search directions of this iteration do not depend on its convergence
        # test (locking acts from the next residual on), so the preconditioner, the
        # product and the Gram are queued before the host reads the norms: one
        # host synchronisation per iteration instead of two (a converged last
        # iteration computes them in vain)
        if precond is not None:
            precond(R, W)
        else:
            W.copy_(R)
        op.apply(W, HW, MW)
        S = [X[c], W] + ([P[c]] if have_p else [])
        HS = [HX[c], HW] + ([HP[c]] if have_p else [])
        MS = [MX[c], MW] + ([MP[c]] if have_p else [])
        m = len(S) * b
        tot = 2 * b + 2 * m * m
        kernels.gram_blocks(S, HS, MS, sym=True, out=xfer[2 * b:tot].view(m, 2 * m))
        xfer_h[:tot].copy_(xfer[:tot], non_blocking=True)
        sync()
        rn_h, mn_h = xfer_h[:b].numpy(), xfer_h[b:2 * b].numpy()
        G = xfer_h[2 * b:tot].numpy().reshape(m, 2 * m)
        # residual relative to the energy scale of the wanted window, so states with
        # eigenvalues close to zero are held to the same absolute accuracy as the rest
        scale = max(np.abs(lam_h[:k]).max(), 1e-300)
        res_h = np.sqrt(np.maximum(rn_h, 0.0)) / (np.maximum(np.abs(lam_h), scale) *
                                                  np.sqrt(np.maximum(mn_h, 1e-300)))
        res = torch.from_numpy(res_h)
        worst = float(res_h[:k].max())
        history.append(worst)
        col_history.append(res_h.copy())
        # lock converged columns from the next residual on (staged with the
        # Rayleigh-Ritz coefficients below)
        mask = (res_h >= tol).astype(np.float64)
        if verbose:
            print(f"  lobpcg[{str(dt)[6:]}] it {it:4d} worst rel. residual {worst:.3e}")
        if worst < tol:
            converged = True
            break
        if stall_iters is not None:
            # stagnation counts close to the target (the first iterations from a
            # random start are not monotone); far from it only a long stall stops
            if worst < min_seen:
                min_seen, min_it = worst, it
            if worst < 100.0 * tol:
                if worst < 0.7 * best:
                    best, best_it = worst, it
                elif it - best_it >= stall_iters:
                    break
            elif it - min_it >= 4 * stall_iters:
                break
        theta, C_h, dropped = _rayleigh_ritz(G, m, b, drop_tol)
        drops.append(dropped)
        CP_h = C_h.copy()
        CP_h[:b] = 0.0
        C, CP, lam_next, colscale = stage_rr.put([C_h, CP_h, theta, mask])
        fresh = bool(refresh) and it % refresh == 0
        kernels.block_update_blocks(S, C, X[1 - c], CP, P[1 - c], c2_first=b)
        if fresh:
            # the images of X are recomputed below; the images of P only need the W
            # and P blocks (the X rows of the P coefficients are zero)
            for src, pimg in ((HS, HP), (MS, MP)):
                kernels.block_update_blocks(src[1:], CP[b:], pimg[1 - c])
        else:
            for src, img, pimg in ((HS, HX, HP), (MS, MX, MP)):
                kernels.block_update_blocks(src, C, img[1 - c], CP, pimg[1 - c], c2_first=b)
        c = 1 - c
        lam, lam_h = lam_next, theta
        have_p = dropped == 0
        if retune is not None and it == retune_iter:
            retune(theta)
        if fresh:
            # recompute the images of X: the updated ones carry the rounding of every
            # update since the last product (FP32 storage)
            op.apply(X[c], HX[c], MX[c])
    sync()
    t_end = time.perf_counter()
    X = X[c]
    timings = {"setup_s": t_iter - t_start, "iterate_s": t_end - t_iter, "restarts": sum(d > 0 for d in drops)}
    out = LOBPCGResult(lam.clone(), X, res, it, history, converged, timings)
    out.col_history = np.array(col_history)
    return out


def _gram_x_first(Gg, m, b):
    """The [W | P | X]-ordered Gram pair of the mixed iteration in [X | W | P]
    order.  Entries of an X row and a W / P column are taken from their transposes
    (W / P row, X column: products with the FP64 images), which also replaces the
    ones that a diagonal tile straddling the W/P and X blocks computed from the FP32
    images."""
    wp = m - b
    out = np.empty_like(Gg)
    for h in (0, 1):
        A = Gg[:, h * m:(h + 1) * m]
        O = out[:, h * m:(h + 1) * m]
        O[:b, :b] = A[wp:, wp:]
        O[b:, b:] = A[:wp, :wp]
        O[b:, :b] = A[:wp, wp:]
        O[:b, b:] = A[:wp, wp:].T
    return out


def lobpcg_mixed(op64, k, b, precond32, X0=None, tol=1e-6, maxiter=400, drop_tol=1e-10, verbose=False,
                 retune=None, retune_at=None, op32=None, seed=0):
    """Block LOBPCG whose iterate X and images HX, MX are FP64 while the search
    directions W, P and their images are FP32 (mixed-precision defect correction).

    Per iteration: the residual R = HX - MX diag(lambda) is formed in FP64 and stored
    in FP32; W = T R by the FP32 multigrid; HW, MW by the FP32 product; the Gram
    [X | W | P]^T [HS | MS] in FP64 from the mixed-type blocks; Rayleigh-Ritz in
    FP64 (host LAPACK); X_new = S C in FP64 and P_new = [W | P] CP in FP32 in one
    pass; the images of P_new by the FP32 product and those of X_new by the FP64
    product (no image is updated from older images).  The FP32 rounding of the corrections perturbs each step
    relative to the correction's own size, which shrinks with the residual, so the
    iteration converges to the FP64 tolerance without the ~1e-3 residual floor of an
    all-FP32 iteration; only X's product is FP64.

    X0: initial (n, >= b) block (FP64 or FP32; default a seeded random block after
    one preconditioner application).  retune(theta): optional callable run
    once, when the worst residual first falls below retune_at (e.g. to move the
    multigrid shift closer to the lowest Ritz value).  Stops when the relative
    residual ||H x - lambda M x|| / (s ||M x||), s = max(|lambda|, max_j<k |lambda_j|),
    of the first k columns is below tol, or after maxiter iterations."""
    with blas_single_thread():
        return _lobpcg_mixed(op64, k, b, precond32, X0, tol, maxiter, drop_tol, verbose, retune, retune_at,
                             op32, seed)


def _lobpcg_mixed(op64, k, b, precond32, X0, tol, maxiter, drop_tol, verbose, retune, retune_at, op32, seed):
    dev = op64.H.device
    n = op64.n
    assert b % 4 == 0 and b >= k and 3 * b <= n, (n, b, k)
    if op32 is None:
        op32 = op64.astype(torch.float32)
    f64, f32 = torch.float64, torch.float32
    t_start = time.perf_counter()

    def blk(dt):
        # written before read, as in _lobpcg
        return torch.empty((n, b), dtype=dt, device=dev)

    X, HX, MX = [blk(f64), blk(f64)], [blk(f64), blk(f64)], [blk(f64), blk(f64)]
    P, HP, MP = [blk(f32), blk(f32)], [blk(f32), blk(f32)], [blk(f32), blk(f32)]
    W, HW, MW, R = blk(f32), blk(f32), blk(f32), blk(f32)
    c = 0
    if X0 is None:
        # seeded random block, smoothed by one preconditioner application
        g = torch.Generator(device=dev).manual_seed(seed)
        R.copy_(torch.rand((n, b), generator=g, dtype=f32, device=dev) - 0.5)
        precond32(R, W)
        X[c].copy_(W)
    else:
        X[c].copy_(X0[:, :b])
    op64.apply(X[c], HX[c], MX[c])
    G = kernels.gram_blocks([X[c]], [HX[c]], [MX[c]], sym=True)
    stage_rr = _Staging(2 * 3 * b * b + 2 * b, dev)
    stage_cs = _Staging(b, dev)
    xfer = torch.empty(2 * b + 18 * b * b, dtype=f64, device=dev)
    xfer_h = torch.empty(xfer.numel(), dtype=f64, pin_memory=dev.type == "cuda")
    theta, C, _ = _rayleigh_ritz(G.cpu().numpy(), b, b, drop_tol)
    C_d, lam = stage_rr.put([C, theta])
    for src in (X, HX, MX):
        kernels.block_update_blocks([src[c]], C_d, src[1 - c])
    c = 1 - c
    lam_h = theta
    have_p = False
    history, col_history, drops = [], [], []
    converged = False
    retuned = retune is None
    it = 0
    res = None
    t_iter = time.perf_counter()
    (colscale,) = stage_cs.put([np.ones(b)])
    have_res = False     # R and its norms already produced by the product of X
    for it in range(1, maxiter + 1):
        if not have_res:
            kernels.residual_norms(HX[c], MX[c], lam, R, colscale, out=xfer[:2 * b])
        precond32(R, W)
        op32.apply(W, HW, MW)
        S = [X[c], W] + ([P[c]] if have_p else [])
        m = len(S) * b
        tot = 2 * b + 2 * m * m
        # Gram over the basis ordered [W | P | X]: its upper tiles then take every
        # X entry from the FP64 images, (HX)^T W rather than X^T (H W) with the FP32
        # product (whose FP32 values of H put an error of ~eps32 ||H|| |X| |W| into
        # the cross terms, a residual floor that does not shrink with W)
        Sg, HSg, MSg = S[1:] + S[:1], [HW] + ([HP[c]] if have_p else []) + [HX[c]], \
            [MW] + ([MP[c]] if have_p else []) + [MX[c]]
        kernels.gram_blocks_typed(Sg, HSg, MSg, sym=True, out=xfer[2 * b:tot].view(m, 2 * m))
        xfer_h[:tot].copy_(xfer[:tot], non_blocking=True)
        sync()
        rn_h, mn_h = xfer_h[:b].numpy(), xfer_h[b:2 * b].numpy()
        G = _gram_x_first(xfer_h[2 * b:tot].numpy().reshape(m, 2 * m), m, b)
        scale = max(np.abs(lam_h[:k]).max(), 1e-300)
        res_h = np.sqrt(np.maximum(rn_h, 0.0)) / (np.maximum(np.abs(lam_h), scale) *
                                                  np.sqrt(np.maximum(mn_h, 1e-300)))
        res = torch.from_numpy(res_h)
        worst = float(res_h[:k].max())
        history.append(worst)
        col_history.append(res_h.copy())
        mask = (res_h >= tol).astype(np.float64)
        if verbose:
            print(f"  lobpcg[mixed] it {it:4d} worst rel. residual {worst:.3e}")
        if worst < tol:
            converged = True
            break
        theta, C_h, dropped = _rayleigh_ritz(G, m, b, drop_tol)
        drops.append(dropped)
        CP_h = C_h.copy()
        CP_h[:b] = 0.0
        C, CP, lam_next, colscale = stage_rr.put([C_h, CP_h, theta, mask])
        # X_new (FP64) and P_new (FP32) in one pass over [X | W | P]
        kernels.block_update_typed(S, C, X[1 - c], CP, P[1 - c], c2_first=b)
        c = 1 - c
        # every image from its own block: P_new is a small difference of larger
        # terms near convergence, so FP32 images updated from those of W and P
        # lose accuracy relative to P_new at every step (the residual then grows
        # below ~1e-4); one FP32 product costs what the two image updates did
        op32.apply(P[c], HP[c], MP[c])
        # the FP64 product of X with the next iteration's residual step fused in
        op64.apply_residual(X[c], HX[c], MX[c], lam_next, R, colscale, xfer[:2 * b])
        have_res = True
        lam, lam_h = lam_next, theta
        have_p = dropped == 0
        if not retuned and retune_at is not None and worst < retune_at:
            retune(theta)
            retuned = True
    sync()
    t_end = time.perf_counter()
    timings = {"setup_s": t_iter - t_start, "iterate_s": t_end - t_iter, "restarts": sum(d > 0 for d in drops)}
    out = LOBPCGResult(lam.clone(), X[c], res, it, history, converged, timings)
    out.col_history = np.array(col_history)
    return out


def solve_mixed(op64, k, b, precond32=None, precond64=None, tol=1e-6, tol32=1e-3, maxiter=400,
                seed=0, verbose=False, drop32=None, stall32=8, refresh32=1, X0=None, b64=None, between=None,
                rr="host"):
    """FP32 LOBPCG to tol32 (or stagnation), then FP64 LOBPCG to tol.  X0: optional
    initial block for the FP32 phase (default: seeded random vectors).

    The FP32 phase recomputes the images HX, MX by SpMM every refresh32 iterations
    (1: every iteration): FP32 rounding of the updated images otherwise makes the
    reported residual up to 30x smaller than the true one.  tol32 = 1e-3 sits just
    above the FP32 floor of the quantum-dot problems (‖H‖/|lambda| ~ 1e3).  b64: block
    of the FP64 phase (default b), started from the lowest b64 FP32 Ritz vectors.
    between: optional callable taking the FP32 result, run before the FP64 phase
    (e.g. to move the preconditioner's shift)."""
    op32 = op64.astype(torch.float32)
    t0 = time.perf_counter()
    r32 = lobpcg(op32, k, b, precond32, X0=X0, tol=tol32, maxiter=maxiter, seed=seed, verbose=verbose,
                 stall_iters=stall32, drop_tol=drop32, refresh=refresh32, rr=rr)
    if between is not None:
        between(r32)
    t1 = time.perf_counter()
    X0 = r32.X.to(torch.float64)
    r64 = lobpcg(op64, k, b64 or b, precond64, X0=X0, tol=tol, maxiter=maxiter, seed=seed, verbose=verbose, rr=rr)
    t2 = time.perf_counter()
    r64.timings.update({"fp32_s": t1 - t0, "fp64_s": t2 - t1, "fp32_iters": r32.iterations,
                        "fp64_iters": r64.iterations, "fp32_restarts": r32.timings["restarts"],
                        "fp64_restarts": r64.timings["restarts"]})
    r64.history = {"fp32": r32.history, "fp64": r64.history}
    r64.col_history = {"fp32": getattr(r32, "col_history", None), "fp64": getattr(r64, "col_history", None)}
    return r64
