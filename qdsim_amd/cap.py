"""Complex-symmetric eigensolver for open systems with a complex absorbing potential.

Problem: (K + M_V - i W) psi = E M psi with real symmetric K, M_V, W and SPD M:
the pencil is complex symmetric (A^T = A, not Hermitian), the eigenvalues are
E = E_r - i Gamma/2 and tau = hbar / (2 |Im E|) (README.md:104-110).  The
reference never solves this problem: it adds |Im CAP| to the real potential,
solves a real eigsh and subtracts a constant i 0.1 cap_strength afterwards
(schrodinger_solver.pyx:160-164, 548-557; fixed_open_system_solver.pyx:321-328).

Method: block preconditioned iteration on S = [X | W | P] with the orthogonal
(Hermitian) Rayleigh-Ritz projection S^H A S y = theta S^H M S y.  The bilinear
projection S^T A S, the Galerkin condition native to complex-symmetric pencils,
produced spurious Ritz values far from the spectrum on the small CAP test case;
the orthogonal one keeps every Ritz value in the field of values of the pencil,
and its fixed points are eigenpairs of the complex-symmetric problem.  The basis
is conditioned with the Hermitian Gram S^H M S (SVQB), and Ritz pairs are selected
by smallest real part (or by distance to a target).  All n-sized work runs in
qdsim_hip kernels on the interleaved (re, im) layout, with X, W, P as separate
contiguous (n, b) blocks:
  * fused CAP SpMM (qds_spmm_cap): (H - iW) X and M X from one pass over the
    pattern with three real value arrays;
  * every complex block is also a real (n, 2m) block, so the MFMA FP64 Gram of
    the real views yields both the bilinear and the Hermitian Gram matrices, and
    the complex update X S C is the real MFMA update with the 2x2-real form of C;
  * the preconditioner is the real lattice multigrid applied to the real and
    imaginary parts (a block of 2b real columns).
complex64 storage for the bulk of the iterations, complex128 refinement.
"""
import numpy as np
import torch

from . import kernels
from .device import blas_single_thread, sync
from .units import EV_TO_J


def _rv(T):
    """Real (n, 2m) view of a complex (n, m) block."""
    return torch.view_as_real(T).reshape(T.shape[0], 2 * T.shape[1])


def _bilinear(G):
    """S^T B from the real Gram of the interleaved views (numpy or torch)."""
    return (G[0::2, 0::2] - G[1::2, 1::2]) + 1j * (G[0::2, 1::2] + G[1::2, 0::2])


def _hermitian(G):
    """S^H B from the real Gram of the interleaved views (numpy or torch)."""
    return (G[0::2, 0::2] + G[1::2, 1::2]) + 1j * (G[0::2, 1::2] - G[1::2, 0::2])


def _real_form(C):
    """2x2-real form of a complex (m, k) matrix acting on interleaved columns."""
    m, k = C.shape
    R = torch.empty((2 * m, 2 * k), dtype=C.real.dtype, device=C.device)
    R[0::2, 0::2] = C.real
    R[1::2, 0::2] = -C.imag
    R[0::2, 1::2] = C.imag
    R[1::2, 1::2] = C.real
    return R


class CAPOperator:
    def __init__(self, indptr, indices, n, H, W, M, cdtype):
        rdt = torch.float32 if cdtype == torch.complex64 else torch.float64
        self.indptr, self.indices, self.n = indptr, indices, int(n)
        self.H = H.to(rdt).contiguous()
        self.W = W.to(rdt).contiguous()
        self.M = M.to(rdt).contiguous()
        self.dtype = cdtype

    def apply(self, X, HX, MX):
        kernels.spmm_cap(self.indptr, self.indices, self.n, self.H, self.W, self.M, X, HX, MX)

    def astype(self, cdtype):
        return CAPOperator(self.indptr, self.indices, self.n, self.H, self.W, self.M, cdtype)


class DeflatedCAPOperator:
    """A' = A + mu (M Q)(M Q)^T with Q^T M Q = I (bilinear, no conjugation): the
    found eigenpairs (lambda_i, q_i) of the complex-symmetric pencil move to
    lambda_i + mu while every other eigenpair is unchanged (its eigenvector is
    bilinearly M-orthogonal to Q).  ``project`` removes the Q components of a
    block, Y <- Y - Q (M Q)^T Y: the complement {x : Q^T M x = 0} is invariant
    under M^-1 A (A^T = A, A Q = M Q Lambda), so the iteration's search directions
    stay there.  The rank-k_q terms run on the interleaved real views: the
    split-row FP64 Gram for (M Q)^T X and the MFMA block update for Q C, in chunks
    of at most 64 columns of Q."""

    _CHUNK = 64

    def __init__(self, base, Q, MQ, mu):
        self.base = base
        self.n = base.n
        self.dtype = base.dtype
        self.H = base.H
        self.mu = float(mu)
        self.Q, self.MQ = [], []
        for j in range(0, Q.shape[1], self._CHUNK):
            q, mq = Q[:, j:j + self._CHUNK], MQ[:, j:j + self._CHUNK]
            if q.shape[1] % 2:      # even widths keep the FP32 real views 16-byte vectors
                z = torch.zeros((q.shape[0], 1), dtype=q.dtype, device=q.device)
                q, mq = torch.cat([q, z], 1), torch.cat([mq, z], 1)
            self.Q.append(q.to(base.dtype).contiguous())
            self.MQ.append(mq.to(base.dtype).contiguous())
        self._tmp = None

    def _low_rank(self, L, R, X, scale, Y):
        """Y += scale * L (R^T X) for chunked (n, k_q) blocks L, R."""
        if self._tmp is None or self._tmp.shape != X.shape:
            self._tmp = torch.empty(X.shape, dtype=self.dtype, device=X.device)
        Xr = _rv(X)
        for l, r in zip(L, R):
            C = _bilinear(kernels.gram(_rv(r), Xr)) * scale
            kernels.block_update(_rv(l), _real_form(C), _rv(self._tmp))
            Y.add_(self._tmp)

    def apply(self, X, HX, MX):
        self.base.apply(X, HX, MX)
        self._low_rank(self.MQ, self.MQ, X, self.mu, HX)

    def project(self, Y):
        self._low_rank(self.Q, self.MQ, Y, -1.0, Y)

    def astype(self, cdtype):
        out = DeflatedCAPOperator.__new__(DeflatedCAPOperator)
        out.base = self.base.astype(cdtype)
        out.n, out.dtype, out.H, out.mu = self.n, cdtype, out.base.H, self.mu
        out.Q = [q.to(cdtype) for q in self.Q]
        out.MQ = [q.to(cdtype) for q in self.MQ]
        out._tmp = None
        return out


def _rr(G, m, nb, drop_tol, target):
    """Orthogonal Rayleigh-Ritz: S^H A S y = theta S^H M S y.  Its Ritz values lie
    in the field of values of the pencil (Re >= lambda_min(H, M), Im <= 0 for W >= 0),
    so no spurious values compete in the selection; at convergence the Ritz pairs
    are eigenpairs of the complex-symmetric pencil."""
    dev = G.device
    g = G.detach().cpu().numpy()
    Ah = _hermitian(g[:, : 2 * m])
    Mh = _hermitian(g[:, 2 * m:])
    Mh = 0.5 * (Mh + Mh.conj().T)
    dg = np.diag(Mh).real.copy()
    ok = dg > 1e-30 * dg.max()
    d = np.where(ok, 1.0 / np.sqrt(np.where(ok, dg, 1.0)), 0.0)
    Mhs = Mh * d[:, None] * d[None, :]
    mu, Q = np.linalg.eigh(Mhs)
    n_ok = int((mu > drop_tol * mu[-1]).sum())
    nk = max(n_ok, nb)
    Bm = (d[:, None] * Q[:, -nk:]) / np.sqrt(np.maximum(mu[-nk:], 1e-300))[None, :]
    Ahat = Bm.conj().T @ Ah @ Bm
    theta, Y = np.linalg.eig(Ahat)
    if target is None:
        order = np.argsort(theta.real)
    else:
        order = np.argsort(np.abs(theta - target))
    sel = order[:nb]
    C = Bm @ Y[:, sel]          # unit Hermitian M-norm columns
    th = theta[sel]
    o2 = np.argsort(th.real)
    return (torch.from_numpy(th[o2].copy()).to(dev), torch.from_numpy(C[:, o2].copy()).to(dev))


def cap_lobpcg(op, k, b, precond=None, X0=None, tol=1e-6, maxiter=500, seed=0, target=None,
               drop_tol=None, stall_iters=None, verbose=False, constraint=None):
    """constraint: optional in-place projection of a complex (n, b) block applied to
    the initial block and to every preconditioned residual (deflation)."""
    # small host eigenproblems: single-threaded BLAS avoids thread wake-up costs
    with blas_single_thread():
        return _cap_lobpcg(op, k, b, precond, X0, tol, maxiter, seed, target, drop_tol, stall_iters, verbose,
                           constraint)


def _cap_lobpcg(op, k, b, precond=None, X0=None, tol=1e-6, maxiter=500, seed=0, target=None,
                drop_tol=None, stall_iters=None, verbose=False, constraint=None):
    dev = op.H.device
    dt = op.dtype
    n = op.n
    assert b % 2 == 0 and 2 * b <= 64 and 3 * b <= n
    if drop_tol is None:
        drop_tol = 1e-5 if dt == torch.complex64 else 1e-10

    def blk():
        return torch.zeros((n, b), dtype=dt, device=dev)

    # contiguous (n, b) complex blocks; X, P and their images double-buffered
    X, HX, MX = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    P, HP, MP = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    W, HW, MW, R = blk(), blk(), blk(), blk()
    c = 0
    if X0 is None:
        g = torch.Generator(device=dev).manual_seed(seed)
        rdt = torch.float32 if dt == torch.complex64 else torch.float64
        X[c].copy_((torch.rand((n, b), generator=g, dtype=rdt, device=dev) - 0.5).to(dt))
        if precond is not None:
            R.copy_(X[c])
            precond(_rv(R), _rv(X[c]))
    else:
        X[c].copy_(X0[:, :b].to(dt))
    if constraint is not None:
        constraint(X[c])
    op.apply(X[c], HX[c], MX[c])
    G = kernels.gram_blocks([_rv(X[c])], [_rv(HX[c])], [_rv(MX[c])])
    lam, C = _rr(G, b, b, drop_tol, target)
    Cr = _real_form(C)
    for src in (X, HX, MX):
        kernels.block_update_blocks([_rv(src[c])], Cr, _rv(src[1 - c]))
    c = 1 - c
    have_p = False
    history = []
    converged = False
    res = None
    best, best_it = float("inf"), 0
    it = 0
    for it in range(1, maxiter + 1):
        kernels.block_residual(HX[c], MX[c], lam.to(torch.complex128).contiguous(), R)
        rn = kernels.coldot(R, R, conj=True).real
        mn = kernels.coldot(MX[c], MX[c], conj=True).real
        # the next search directions and their Gram do not depend on the
        # convergence test: queue them before the host reads the residuals
        if precond is not None:
            precond(_rv(R), _rv(W))
        else:
            W.copy_(R)
        if constraint is not None:
            constraint(W)
        op.apply(W, HW, MW)
        S = [X[c], W] + ([P[c]] if have_p else [])
        HS = [HX[c], HW] + ([HP[c]] if have_p else [])
        MS = [MX[c], MW] + ([MP[c]] if have_p else [])
        m = len(S) * b
        G = kernels.gram_blocks([_rv(x) for x in S], [_rv(x) for x in HS], [_rv(x) for x in MS])
        scale = lam[:k].abs().max().clamp_min(1e-300)
        res = rn.clamp_min(0).sqrt() / (torch.maximum(lam.abs(), scale) * mn.clamp_min(1e-300).sqrt())
        worst = float(res[:k].max().item())
        history.append(worst)
        if verbose:
            print(f"  cap[{str(dt)[6:]}] it {it:4d} worst rel. residual {worst:.3e}")
        if worst < tol:
            converged = True
            break
        if stall_iters is not None:
            if worst < 0.7 * best:
                best, best_it = worst, it
            elif it - best_it >= stall_iters:
                break
        lam, C = _rr(G, m, b, drop_tol, target)
        CP = C.clone()
        CP[:b] = 0
        Cr, CPr = _real_form(C), _real_form(CP)
        for src, img, pimg in ((S, X, P), (HS, HX, HP), (MS, MX, MP)):
            kernels.block_update_blocks([_rv(x) for x in src], Cr, _rv(img[1 - c]), CPr, _rv(pimg[1 - c]),
                                        c2_first=2 * b)
        c = 1 - c
        have_p = True
    sync()
    return lam, X[c], res, it, converged, history


def _cap_pair(op64, k, b, pre32, pre64, tol, tol32, maxiter, seed, verbose):
    """FP32 then FP64 block iteration for the k lowest-real-part eigenpairs."""
    op32 = op64.astype(torch.complex64)
    lam32, X32, _, it32, _, h32 = cap_lobpcg(op32, k, b, pre32, tol=tol32, maxiter=maxiter, seed=seed,
                                            stall_iters=10, verbose=verbose,
                                            constraint=getattr(op32, "project", None))
    lam, X, res, it64, conv, h64 = cap_lobpcg(op64, k, b, pre64, X0=X32.to(torch.complex128), tol=tol,
                                             maxiter=maxiter, seed=seed, verbose=verbose,
                                             constraint=getattr(op64, "project", None))
    return lam, X, res, it32, it64, conv, h32, h64


def _csym_chol(G):
    """Lower-triangular L with G = L L^T (transpose, no conjugation) for a complex
    symmetric G near the identity (unpivoted elimination)."""
    n = G.shape[0]
    L = np.zeros_like(G)
    for j in range(n):
        L[j, j] = np.sqrt(G[j, j] - np.sum(L[j, :j] ** 2))
        L[j + 1:, j] = (G[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    return L


def _deflation_block(op64, op, X):
    """Bilinearly M-orthonormal basis of span(X) in the complement of the current
    deflation basis (op.project when op deflates), and its M image."""
    Q = X.clone()
    if hasattr(op, "project"):
        op.project(Q)
    HQ, MQ = torch.empty_like(Q), torch.empty_like(Q)
    op64.apply(Q, HQ, MQ)
    G = _bilinear(kernels.gram(_rv(Q), _rv(MQ))).cpu().numpy()
    G = 0.5 * (G + G.T)
    Linv_t = np.linalg.inv(_csym_chol(G)).T
    Cr = _real_form(torch.from_numpy(Linv_t).to(Q.device))
    Qo, MQo = torch.empty_like(Q), torch.empty_like(Q)
    kernels.block_update(_rv(Q), Cr, _rv(Qo))
    kernels.block_update(_rv(MQ), Cr, _rv(MQo))
    return Qo, MQo


def _bilinear_pairs(op64, X):
    """Bilinear Rayleigh quotients and images of the columns of X (original pencil)."""
    Xk = X.to(torch.complex128).contiguous()
    MX = torch.empty_like(Xk)
    HX = torch.empty_like(Xk)
    op64.apply(Xk, HX, MX)
    bil = (Xk * MX).sum(0)
    return (Xk * HX).sum(0) / bil, Xk, MX, bil


def solve_cap(prob, k, b=None, target=None, tol=1e-6, tol32=3e-4, maxiter=600, seed=0, verbose=False):
    """Lowest-real-part (or nearest-target, eV) eigenpairs of the CAP pencil of a
    SchrodingerProblem whose CAP coefficient is set.  Returns (E complex eV (np),
    eigenvectors (n_full, k) complex128 np, Hermitian M_nm-normalised).

    Nearest-target mode (the states nearest an energy, as shift-invert eigs(H, k, M, sigma) selects them) is
    computed from the bottom of the spectrum: batches of the lowest-real-part
    eigenpairs, each batch found on the pencil with the previous ones deflated
    (DeflatedCAPOperator), until the computed set brackets the k eigenvalues
    nearest the target (every eigenvalue not yet computed has a real part above
    the largest computed one, hence is farther from the target than the k-th
    nearest).  Selecting Ritz pairs by distance to an interior target inside the
    block iteration converged to spurious Ritz values on the full-size device."""
    if prob.H is None or prob.W is None:
        prob.assemble()
    if b is None:
        b = min(32, max(8, ((k + 8) + 1) // 2 * 2))
    H = kernels.csr_take(prob.r_src, prob.H)
    W = kernels.csr_take(prob.r_src, prob.W)
    M = kernels.csr_take(prob.r_src, prob.M)
    op64 = CAPOperator(prob.r_indptr, prob.r_indices, prob.n_free, H, W, M, torch.complex128)
    pre32 = prob.preconditioner(2 * b, torch.float32)

    def pre64(Rr, Wr, _buf={}):
        if pre32.takes_fp64 and Rr.stride(0) == Rr.shape[1] and Wr.stride(0) == Wr.shape[1]:
            pre32(Rr, Wr)
            return
        if "r" not in _buf:
            _buf["r"] = torch.empty(Rr.shape, dtype=torch.float32, device=Rr.device)
            _buf["w"] = torch.empty(Rr.shape, dtype=torch.float32, device=Rr.device)
        _buf["r"].copy_(Rr)
        pre32(_buf["r"], _buf["w"])
        Wr.copy_(_buf["w"])

    if target is None:
        lam, X, res, it32, it64, conv, h32, h64 = _cap_pair(op64, k, b, pre32, pre64, tol, tol32, maxiter,
                                                             seed, verbose)
        sel = torch.arange(k, device=lam.device)
        X = X[:, sel]
        res = res[sel] if res is not None else None
    else:
        kb = max(1, b - 8)                      # eigenpairs per batch
        Qs, lams, ress = [], [], []
        it32 = it64 = 0
        conv = True
        h32, h64 = [], []
        op = op64
        Qd = MQd = None             # deflation basis, Q^T M Q = I (bilinear)
        for batch in range(16):
            lb, Xb, rb, i32, i64, cb, g32, g64 = _cap_pair(op, kb, b, pre32, pre64, tol, tol32, maxiter,
                                                            seed + batch, verbose)
            it32, it64, conv = it32 + i32, it64 + i64, conv and cb
            h32, h64 = h32 + g32, h64 + g64
            lb_o, Xb_o, MXb, bil = _bilinear_pairs(op64, Xb[:, :kb])
            Qs.append(Xb_o / bil.sqrt()[None, :])
            lams.append(lb_o)
            ress.append(rb[:kb] if rb is not None else torch.zeros(kb, dtype=torch.float64, device=lb.device))
            lam_all = torch.cat(lams)
            dist = (lam_all - target).abs()
            if len(lam_all) >= k:
                dk = torch.sort(dist).values[k - 1]
                if float(lam_all.real.max() - target) >= float(dk):
                    break
            Qn, MQn = _deflation_block(op64, op, Qs[-1])
            Qd = Qn if Qd is None else torch.cat([Qd, Qn], 1)
            MQd = MQn if MQd is None else torch.cat([MQd, MQn], 1)
            # shift the found eigenvalues well above the window searched next and
            # keep the search directions bilinearly M-orthogonal to them
            mu = 4.0 * float(lam_all.abs().max()) + 4.0 * abs(float(target)) + 1.0
            op = DeflatedCAPOperator(op64, Qd, MQd, mu)
        else:
            raise RuntimeError("solve_cap: the target window was not bracketed after 16 batches")
        lam = torch.cat(lams)
        X = torch.cat(Qs, 1)
        res = torch.cat(ress)
        sel = torch.argsort((lam - target).abs())[:k]
        sel = sel[torch.argsort(lam[sel].real)]
        X = X[:, sel]
        res = res[sel]
    lam, Xk, MX, bil = _bilinear_pairs(op64, X)
    # bilinear Rayleigh quotient x^T A x / x^T M x: second-order accurate for a
    # complex-symmetric pencil (the Hermitian Ritz value of the orthogonal
    # projection is only first-order accurate for non-Hermitian A)
    # Hermitian M-normalisation with the phase making psi^T M psi real positive
    Xk = Xk * torch.exp(-0.5j * torch.angle(bil))[None, :]
    herm = (Xk.conj() * MX * torch.exp(-0.5j * torch.angle(bil))[None, :]).sum(0).real
    Xk = Xk / herm.clamp_min(1e-300).sqrt()[None, :]
    full = torch.zeros((prob.space.n, k), dtype=torch.complex128, device=Xk.device)
    full[prob.old_of_new] = Xk
    prob.last_cap_info = {"fp32_iters": it32, "fp64_iters": it64, "converged": conv,
                          "residuals": res.cpu().numpy() if res is not None else None,
                          "history": {"c64": h32, "c128": h64}}
    return lam.cpu().numpy(), full.cpu().numpy()
