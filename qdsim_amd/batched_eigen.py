"""Batched block eigensolver: many problems on one mesh in one launch per stage.

No reference counterpart: the reference solves one Hamiltonian per call
(qdsim_cython/solvers/eigen_solver.pyx:225-266 and the Schroedinger solvers
around it), so a parameter sweep is a loop of solves.  Here a batch of potentials
on one structured mesh (the gate-parameter sets of an inverse-design loop) shares
the mesh, the CSR pattern and lattice stencil, the mass matrix and the kinetic
part of the multigrid hierarchy; each problem keeps its own H values, blocks,
multigrid diagonals, shift and Rayleigh-Ritz.  Every stage of the block iteration
is one launch for the whole batch:

  residual norms (qds_residual_norms_batched) -> multigrid V-cycles
  (qds_mg_vcycle_batched) -> H/M product (qds_lattice_spmm_batched) -> Gram
  (qds_gram_blocks_batched) -> Rayleigh-Ritz, residual test, locking and freeze
  (qds_rayleigh_ritz, one workgroup per problem) -> block updates
  (qds_block_update_blocks_batched)

and the host reads the batch's status one iteration behind, as the single-problem
device path does.  A problem whose residual test passes keeps its block (freeze)
while the others continue; the iteration ends when every problem has converged
(or stalled, in the FP32 phase).
"""
import time

import numpy as np
import torch

from . import kernels
from .device import sync
from .lobpcg import _StallTest, _default_drop_tol
from .schrodinger import SchrodingerProblem, default_block, shift_below


class BatchedLatticeOperator:
    """H_i (batch, slots) and a shared M (slots,) in the lattice form."""

    def __init__(self, stencil, n, H, M, dtype):
        self.stencil = stencil
        self.n = int(n)
        self.H = H.to(dtype).contiguous()
        self.M = M.to(dtype).contiguous()
        self.batch = int(H.shape[0])
        self.dtype = dtype

    def apply(self, X, HX, MX):
        kernels.lattice_spmm_auto(self.stencil, self.H, X, HX, self.M, MX)

    def astype(self, dtype):
        return BatchedLatticeOperator(self.stencil, self.n, self.H, self.M, dtype)


class BatchedResult:
    def __init__(self, lam, X, res, iters, history, converged, timings):
        self.eigenvalues = lam          # (batch, b) device FP64
        self.X = X                      # (batch, n, b) device
        self.residuals = res            # (batch, b) numpy
        self.iterations = iters
        self.history = history          # per iteration: worst residual of every problem
        self.converged = converged      # (batch,) bool numpy
        self.timings = timings


def lobpcg_batched(op, k, b, precond=None, X0=None, tol=1e-6, maxiter=400, drop_tol=None, seed=0,
                   stall_iters=None, refresh=None, verbose=False, graph=True):
    """Lowest k eigenpairs of each (H_i, M) of a BatchedLatticeOperator, block size b
    (b % 4 == 0, k <= b, 3 b <= 96, 3 b <= n).  precond(R, W): (batch, n, b) blocks.
    Same iteration, stopping rule and soft locking as the single-problem LOBPCG;
    refresh (truthy) recomputes the images of X by product every iteration.

    graph: after two eager iterations, the iteration (all its launches and the
    status copy) is captured in a HIP graph, one per parity of the double buffers,
    and replayed: one graph launch per iteration instead of ~15 kernel launches
    from Python.  The captured work is the eager work (same kernels, same buffers;
    no memsets or allocations inside), so the trajectory is the same."""
    dev = op.H.device
    dt = op.dtype
    n, B = op.n, op.batch
    assert b % 4 == 0 and b >= k and 3 * b <= n and 3 * b <= 96, (n, b, k)
    if drop_tol is None:
        drop_tol = _default_drop_tol(dt)
    t_start = time.perf_counter()

    def blk():
        return torch.zeros((B, n, b), dtype=dt, device=dev)

    f64 = dict(dtype=torch.float64, device=dev)
    X, HX, MX = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    P, HP, MP = [blk(), blk()], [blk(), blk()], [blk(), blk()]
    W, HW, MW, R = blk(), blk(), blk(), blk()
    m = 3 * b
    nst = kernels.rr_status_len(b)
    G = torch.zeros((B, m, 2 * m), **f64)
    norms = torch.zeros((B, 2 * b), **f64)
    lam = torch.zeros((B, b), **f64)
    C = torch.zeros((B, m, b), **f64)
    CP = torch.zeros((B, m, b), **f64)
    colscale = torch.ones((B, b), **f64)
    status = torch.zeros((B, nst), **f64)
    status_h = torch.zeros((2, B, nst), dtype=torch.float64, pin_memory=dev.type == "cuda")
    events = [torch.cuda.Event(), torch.cuda.Event()]
    c = 0
    if X0 is None:
        g = torch.Generator(device=dev).manual_seed(seed)
        X[c].copy_(torch.rand((B, n, b), generator=g, dtype=dt, device=dev) - 0.5)
        if precond is not None:
            R.copy_(X[c])
            precond(R, X[c])
    else:
        X[c].copy_(X0[:, :, :b].to(dt))
    op.apply(X[c], HX[c], MX[c])
    G0 = torch.zeros((B, b, 2 * b), **f64)
    C0 = torch.zeros((B, b, b), **f64)
    kernels.gram_blocks_batched([X[c]], [HX[c]], [MX[c]], sym=True, out=G0)
    kernels.rayleigh_ritz(G0, lam, C0, status, k, tol, drop_tol)
    for src in (X, HX, MX):
        kernels.block_update_blocks_batched([src[c]], C0, src[1 - c])
    c = 1 - c
    CPb = CP[:, b:]

    def iterate(c):
        kernels.residual_norms_batched(HX[c], MX[c], lam, R, colscale, norms)
        if precond is not None:
            precond(R, W)
        else:
            W.copy_(R)
        op.apply(W, HW, MW)
        S, HS, MS = [X[c], W, P[c]], [HX[c], HW, HP[c]], [MX[c], MW, MP[c]]
        kernels.gram_blocks_batched(S, HS, MS, sym=True, out=G)
        kernels.rayleigh_ritz(G, lam, C, status, k, tol, drop_tol, norms=norms, CP=CP, colscale=colscale,
                              freeze=True)
        kernels.block_update_blocks_batched(S, C, X[1 - c], CP, P[1 - c], c2_first=b)
        if refresh:
            for src, pimg in ((HS, HP), (MS, MP)):
                kernels.block_update_blocks_batched(src[1:], CPb, pimg[1 - c])
            op.apply(X[1 - c], HX[1 - c], MX[1 - c])
        else:
            for src, img, pimg in ((HS, HX, HP), (MS, MX, MP)):
                kernels.block_update_blocks_batched(src, C, img[1 - c], CP, pimg[1 - c], c2_first=b)

    graphs = {}

    def launch(it, c):
        slot = it % 2
        if graph and it >= 3:
            g = graphs.get((c, slot))
            if g is None:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    iterate(c)
                    status_h[slot].copy_(status, non_blocking=True)
                graphs[(c, slot)] = g
            g.replay()
        else:
            iterate(c)
            status_h[slot].copy_(status, non_blocking=True)
        events[slot].record()

    stall = [_StallTest(tol, stall_iters) for _ in range(B)]
    done = np.zeros(B, dtype=bool)
    conv = np.zeros(B, dtype=bool)
    history, res = [], None
    t_iter = time.perf_counter()
    launch(1, c)
    c = 1 - c
    it = 0
    for it in range(1, maxiter + 1):
        if it < maxiter:
            launch(it + 1, c)
            c = 1 - c
        events[it % 2].synchronize()
        st = status_h[it % 2].numpy().copy()
        worst = st[:, 0].copy()
        history.append(worst)
        res = st[:, 8:8 + b]
        conv = st[:, 1] > 0
        for i in range(B):
            if not done[i] and (conv[i] or stall[i].stalled(it, float(worst[i]))):
                done[i] = True
        done |= conv
        if verbose:
            print(f"  batched lobpcg[{str(dt)[6:]}] it {it:4d} worst {worst.max():.3e} "
                  f"converged {int(conv.sum())}/{B}")
        if done.all():
            break
    sync()
    t_end = time.perf_counter()
    timings = {"setup_s": t_iter - t_start, "iterate_s": t_end - t_iter, "graphs": len(graphs)}
    out = BatchedResult(lam.clone(), X[c].clone(), res.copy(), it, history, conv.copy(), timings)
    del graphs
    return out


class BatchedSchrodinger:
    """A batch of closed-system problems on one structured mesh with one effective
    mass and per-problem potentials (eV at the quadrature points), solved together:
    H_i = K + M(V_i), M shared.  Mixed precision as the single-problem solve: FP32
    blocks and products with FP64 Gram and Rayleigh-Ritz to tol32, then FP64 to
    tol; the preconditioner is the batched FP32 multigrid with per-problem shifts
    below each problem's lowest eigenvalue (moved from the coarse and the FP32 Ritz
    values as in SchrodingerProblem.solve); the start is the batch's own solution on
    a mesh with 4x fewer cells per direction (nested iteration) when the mesh allows
    it, else seeded random vectors."""

    def __init__(self, mesh, m_star_func, nested=True):
        self.mesh = mesh
        self.prob = SchrodingerProblem(mesh, m_star_func, None)
        p = self.prob
        self.space = p.space
        self.c_q = p.c_q
        # shared mass values in lattice slots (FP64)
        _, M_full, _ = p.space.assemble(c=p.c_q, v=None, stiffness_mass=False, mass=True)
        self.M_full = M_full
        self.M_slots = kernels.csr_take(p.s_src, M_full, dtype=torch.float64)
        self.m_star_func = m_star_func
        self.nested = nested
        self._coarse = None

    def quad_points_nm(self):
        return self.space.quad_points()

    def _operator(self, v_q):
        """FP64 BatchedLatticeOperator for potentials v_q (batch, E, Q) eV."""
        p = self.prob
        B = v_q.shape[0]
        H = torch.empty((B, p.s_src.numel()), dtype=torch.float64, device=v_q.device)
        for i in range(B):
            h_full, _, _ = self.space.assemble(c=self.c_q, v=v_q[i].contiguous(), stiffness_mass=True, mass=False)
            H[i] = kernels.csr_take(p.s_src, h_full, dtype=torch.float64)
        return BatchedLatticeOperator(p.stencil, p.n_free, H, self.M_slots, torch.float64)

    def _preconditioner(self, v_q, b):
        p = self.prob
        c_e = self.c_q.mean(dim=1) if self.c_q.dim() == 2 else self.c_q
        v_e = v_q.mean(dim=2)
        vmin = v_e.min(dim=1).values
        vmax = v_e.max(dim=1).values
        floor = (vmin - 0.05 * (vmax - vmin) - 1e-3).cpu().numpy()
        from .lattice import BatchedLatticeMG
        return BatchedLatticeMG(p.lat, c_e, v_e, floor, b, dtype=torch.float32), floor

    def _nested_start(self, vfun, k, b, factor=4):
        cx, cy = self.mesh.nx - 1, self.mesh.ny - 1
        if not self.nested or cx % factor or cy % factor or min(cx, cy) // factor < 32:
            return None, None
        if self._coarse is None:
            from .interpolator import FEInterpolator
            from .mesh import SimpleMesh
            cm = SimpleMesh(cx // factor + 1, cy // factor + 1, self.mesh.Lx, self.mesh.Ly, self.mesh.element_order)
            cb = BatchedSchrodinger(cm, self.m_star_func, nested=False)
            L = self.prob.lat
            iy, ix = np.divmod(np.arange(L.n_free), L.nx_i)
            dev = self.space.dev
            px = torch.from_numpy((ix + 1) * (self.mesh.Lx / (L.NX - 1))).to(dev)
            py = torch.from_numpy((iy + 1) * (self.mesh.Ly / (L.NY - 1))).to(dev)
            self._coarse = (cb, FEInterpolator(cm), px, py)
        cb, interp, px, py = self._coarse
        xq, yq = cb.quad_points_nm()
        v_c = vfun(xq, yq)
        op = cb._operator(v_c).astype(torch.float32)
        pre, _ = cb._preconditioner(v_c, b)
        r = lobpcg_batched(op, k, b, pre, tol=1e-2, maxiter=200, stall_iters=8, refresh=1)
        B = v_c.shape[0]
        full = torch.zeros((cb.space.n, B, b), dtype=torch.float64, device=r.X.device)
        full[cb.prob.old_of_new] = r.X.double().permute(1, 0, 2)
        vals = interp.interpolate_points(px, py, full.reshape(cb.space.n, B * b))
        X0 = vals.reshape(-1, B, b).permute(1, 0, 2).contiguous()
        self.nested_iters = r.iterations
        return X0, r.eigenvalues

    def solve(self, vfun, k, b=None, tol=1e-6, tol32=3e-3, maxiter=400, seed=0, verbose=False):
        """vfun(xq, yq) -> (batch, E, Q) potentials (eV) at quadrature points given in
        nm.  Returns (eigenvalues (batch, k) numpy eV, full-numbering eigenvectors
        (batch, n, k) FP64 device M_nm-normalised, info)."""
        t0 = time.perf_counter()
        p = self.prob
        if b is None:
            b = default_block(k, p.n_free)
        xq, yq = self.quad_points_nm()
        v_q = vfun(xq, yq)
        B = v_q.shape[0]
        op64 = self._operator(v_q)
        op32 = op64.astype(torch.float32)
        pre, floor = self._preconditioner(v_q, b)
        t_n = time.perf_counter()
        X0, theta = self._nested_start(vfun, k, b)
        if theta is not None:
            th = theta.cpu().numpy()
            pre.set_sigma([shift_below(th[i], floor[i], 0.25) for i in range(B)])
        t_32 = time.perf_counter()
        r32 = lobpcg_batched(op32, k, b, pre, X0=X0, tol=tol32, maxiter=maxiter, seed=seed, stall_iters=8,
                             refresh=1, verbose=verbose)
        th = r32.eigenvalues.cpu().numpy()
        pre.set_sigma([shift_below(th[i], floor[i], 0.1) for i in range(B)])

        def pre64(R, W, _buf={}):
            if pre.takes_fp64:
                pre(R, W)
                return
            if "r" not in _buf:
                _buf["r"] = torch.empty(R.shape, dtype=torch.float32, device=R.device)
                _buf["w"] = torch.empty(R.shape, dtype=torch.float32, device=R.device)
            _buf["r"].copy_(R)
            pre(_buf["r"], _buf["w"])
            W.copy_(_buf["w"])
        t_64 = time.perf_counter()
        r64 = lobpcg_batched(op64, k, b, pre64, X0=r32.X.double(), tol=tol, maxiter=maxiter, seed=seed,
                             verbose=verbose)
        lam = r64.eigenvalues[:, :k]
        X = r64.X[:, :, :k]
        full = torch.zeros((B, self.space.n, k), dtype=torch.float64, device=X.device)
        full[:, p.old_of_new] = X.double()
        sync()
        t1 = time.perf_counter()
        info = {"converged": r64.converged, "residuals": r64.residuals[:, :k], "block": b,
                "fp32_iters": r32.iterations, "fp64_iters": r64.iterations,
                "nested_iters": getattr(self, "nested_iters", 0) if theta is not None else 0,
                "timings": {"setup_s": t_n - t0, "nested_s": t_32 - t_n, "fp32_s": t_64 - t_32,
                            "fp64_s": t1 - t_64, "solve_s": t1 - t0}}
        return lam.cpu().numpy(), full, info
