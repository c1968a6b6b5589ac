"""Lattice view of a structured FE mesh and the multigrid block preconditioner.

Every DOF of an order-p SimpleMesh sits on the (p*cx+1) x (p*cy+1) lattice of the
p-times refined grid (cx, cy cells).  The eigensolver numbers the free
(interior) DOFs in lattice row-major order, so a block vector is a plain
(ny_i * nx_i, b) array and the preconditioner can work on 2-D stencils.

Preconditioner: one multigrid V-cycle (qds_mg_vcycle) for the low-order-refined
operator K_lor + (V - sigma) M_lump on the lattice: five-point P1 stiffness of the
refined right-triangle mesh (edge weight c * h_perp / h_par, c = hbar^2/2m*
averaged over the FE elements touching the edge), lumped P1 mass h_x h_y, shift
sigma below min(V) so the operator is SPD.  Coarse levels are rediscretised:
coarse edge weight = harmonic mean of the two fine edges on it, coarse mass term
= R m with R = P^T of bilinear interpolation.
"""
import ctypes

import numpy as np
import torch

from . import _lib, kernels
from .device import device


class Lattice:
    """Lattice geometry of a SimpleMesh (lengths in nm)."""

    def __init__(self, mesh):
        self.p = int(getattr(mesh, "element_order", 1))
        self.cx, self.cy = mesh.nx - 1, mesh.ny - 1
        self.NX = self.p * self.cx + 1
        self.NY = self.p * self.cy + 1
        self.nx_i, self.ny_i = self.NX - 2, self.NY - 2
        self.Lx = mesh.Lx * 1e9
        self.Ly = mesh.Ly * 1e9
        self.hx = self.Lx / (self.p * self.cx)
        self.hy = self.Ly / (self.p * self.cy)
        self.dxc = self.Lx / self.cx
        self.dyc = self.Ly / self.cy
        lx = np.asarray(mesh.lattice_x)
        ly = np.asarray(mesh.lattice_y)
        conn = mesh.get_elements()
        used = np.zeros(mesh.num_nodes, dtype=bool)
        used[conn.ravel()] = True
        interior = (lx > 0) & (lx < self.NX - 1) & (ly > 0) & (ly < self.NY - 1) & used
        new = np.full(mesh.num_nodes, -1, dtype=np.int64)
        new[interior] = (ly[interior] - 1) * self.nx_i + (lx[interior] - 1)
        self.new_of_old = new
        self.n_free = self.nx_i * self.ny_i
        if interior.sum() != self.n_free:
            raise ValueError("mesh DOFs do not cover the interior lattice one-to-one")
        self.free_dofs = np.nonzero(interior)[0]
        # old id of each lattice-ordered free DOF
        old = np.empty(self.n_free, dtype=np.int64)
        old[new[interior]] = np.nonzero(interior)[0]
        self.old_of_new = old

    def element_at(self, x, y):
        """FE element index containing points (x, y) in nm (device tensors)."""
        fx = x / self.dxc
        fy = y / self.dyc
        i = torch.clamp(torch.floor(fx), 0, self.cx - 1)
        j = torch.clamp(torch.floor(fy), 0, self.cy - 1)
        u = fx - i
        v = fy - j
        upper = (u + v > 1.0).to(torch.int64)
        return 2 * (j.to(torch.int64) * self.cx + i.to(torch.int64)) + upper


class LatticeMG:
    """Block multigrid preconditioner W = V-cycle(R) on the interior lattice."""

    # V(1,1) with 10 coarse sweeps: on the quantum-dot problems the block iteration
    # counts match V(2,2) with 40 coarse sweeps (the spectrum, not the preconditioner
    # quality, limits convergence) at half the multigrid cost
    def __init__(self, lattice, c_elem, v_elem, sigma, b, dtype=torch.float32, nu_pre=1, nu_post=1,
                 coarse_sweeps=10, omega=0.8, min_coarse=3):
        self.lat = lattice
        self.b = int(b)
        self.dtype = dtype
        self.nu_pre, self.nu_post, self.coarse_sweeps, self.omega = nu_pre, nu_post, coarse_sweeps, omega
        dev = device()
        L = lattice
        c_elem = c_elem.to(dev, torch.float64).contiguous()
        v_elem = v_elem.to(dev, torch.float64).contiguous()
        f64 = dict(dtype=torch.float64, device=dev)
        self.sigma = float(sigma)
        # fine full-lattice edge weights, mass term and lumped mass (one kernel), then
        # per level the unknowns' diagonals and weights and the next coarser lattice
        NX, NY = L.NX, L.NY
        WE = torch.empty((NY, NX - 1), **f64)
        WN = torch.empty((NY - 1, NX), **f64)
        Mterm = torch.empty((NY, NX), **f64)
        Mass = torch.empty((NY, NX), **f64)
        kernels.mg_lattice_coefficients(NX, NY, L.cx, L.cy, L.hx, L.hy, 1.0 / L.dxc, 1.0 / L.dyc, c_elem, v_elem,
                                        self.sigma, WE, WN, Mterm, Mass)
        shapes = []
        while True:
            shapes.append((NX, NY))
            nxi, nyi = NX - 2, NY - 2
            if not (nxi % 2 == 1 and nyi % 2 == 1 and (nxi - 1) // 2 >= min_coarse
                    and (nyi - 1) // 2 >= min_coarse):
                break
            NX, NY = (NX + 1) // 2, (NY + 1) // 2
        self.n_levels = len(shapes)
        # every level's u, f, t in one zeroed buffer
        sizes = [-(-(nx - 2) * (ny - 2) * self.b // 64) * 64 for nx, ny in shapes]   # 256-byte aligned slices
        work = torch.zeros(3 * sum(sizes), dtype=dtype, device=dev)
        self.levels_data = []
        off = 0
        for l, (NX, NY) in enumerate(shapes):
            n = (NX - 2) * (NY - 2)
            lv = {"nx": NX - 2, "ny": NY - 2,
                  "d": torch.empty(n, dtype=dtype, device=dev), "we": torch.empty(n, dtype=dtype, device=dev),
                  "wn": torch.empty(n, dtype=dtype, device=dev), "d64": torch.empty((NY - 2, NX - 2), **f64),
                  "mass": torch.empty((NY - 2, NX - 2), **f64)}
            kernels.mg_lattice_interior(dtype, NX, NY, WE, WN, Mterm, Mass, lv["d"], lv["we"], lv["wn"],
                                        lv["d64"], lv["mass"])
            for name in ("u", "f", "t"):
                lv[name] = work[off:off + n * self.b].view(n, self.b)
                off += sizes[l]
            self.levels_data.append(lv)
            if l + 1 < len(shapes):
                NXc, NYc = shapes[l + 1]
                WEc = torch.empty((NYc, NXc - 1), **f64)
                WNc = torch.empty((NYc - 1, NXc), **f64)
                Mc = torch.empty((NYc, NXc), **f64)
                Massc = torch.empty((NYc, NXc), **f64)
                kernels.mg_lattice_coarsen(NX, NY, WE, WN, Mterm, Mass, WEc, WNc, Mc, Massc)
                WE, WN, Mterm, Mass = WEc, WNc, Mc, Massc
        self._structs = (_lib.MGLevel * self.n_levels)()
        for l, lv in enumerate(self.levels_data):
            s = self._structs[l]
            s.nx, s.ny = lv["nx"], lv["ny"]
            s.d, s.we, s.wn = lv["d"].data_ptr(), lv["we"].data_ptr(), lv["wn"].data_ptr()
            s.u, s.f, s.t = lv["u"].data_ptr(), lv["f"].data_ptr(), lv["t"].data_ptr()
            s.ldu = s.ldf = 0

    def set_sigma(self, sigma):
        """Move the shift of every level to sigma (diagonals updated in place:
        d = d(sigma_0) - (sigma - sigma_0) * lumped mass)."""
        sigma = float(sigma)
        if sigma == self.sigma:
            return
        for lv in self.levels_data:
            lv["d64"] = lv["d64"] - (sigma - self.sigma) * lv["mass"]
            lv["d"].copy_(lv["d64"].reshape(-1).to(self.dtype))
        self.sigma = sigma

    @property
    def takes_fp64(self):
        """An FP32 hierarchy whose fused V(1,1) keeps level 0 out of the
        single-workgroup coarse cycle (level 0 above 24576 work items of 16 bytes
        or 4 bytes per lane) can run on FP64 blocks directly."""
        if self.dtype != torch.float32 or self.nu_pre != 1 or self.nu_post != 1 or self.n_levels < 2:
            return False
        lv = self.levels_data[0]
        vec = 4 if self.b % 4 == 0 else 1
        return lv["nx"] * lv["ny"] * (self.b // vec) > 24576

    @staticmethod
    def _coarsen(WE, WN, Mterm, NX, NY):
        NXc, NYc = (NX + 1) // 2, (NY + 1) // 2

        def hmean(a, b):
            return 2.0 * a * b / (a + b)

        WEc = hmean(WE[0::2, 0::2], WE[0::2, 1::2])
        WNc = hmean(WN[0::2, 0::2], WN[1::2, 0::2])
        # R m with bilinear weights; pad with zeros outside
        Mp = torch.nn.functional.pad(Mterm[None, None], (1, 1, 1, 1))[0, 0]
        w = torch.tensor([0.5, 1.0, 0.5], dtype=Mterm.dtype, device=Mterm.device)
        Mc = torch.zeros((NYc, NXc), dtype=Mterm.dtype, device=Mterm.device)
        for a in range(3):
            for c in range(3):
                Mc += w[a] * w[c] * Mp[a:a + 2 * NYc - 1:2, c:c + 2 * NXc - 1:2]
        return WEc, WNc, Mc, NXc, NYc

    def __call__(self, R, W):
        """W = V-cycle(R); R dense (n, b), W may be a column slice (row stride >= b).
        An FP32 hierarchy also takes FP64 R and W (qds_mg_vcycle_io) when its cycle
        allows it (see ``takes_fp64``)."""
        assert R.dtype == W.dtype
        if R.dtype != self.dtype:
            assert self.dtype == torch.float32 and R.dtype == torch.float64 and self.takes_fp64
            s0 = self._structs[0]
            f_old, u_old, ldf_old, ldu_old = s0.f, s0.u, s0.ldf, s0.ldu
            s0.f, s0.ldf = R.data_ptr(), R.stride(0)
            s0.u, s0.ldu = W.data_ptr(), W.stride(0)
            try:
                kernels.mg_vcycle_io(_lib.F32, _lib.F64, self._structs, self.n_levels, self.b, self.nu_pre,
                                     self.nu_post, self.coarse_sweeps, self.omega)
            finally:
                s0.f, s0.u, s0.ldf, s0.ldu = f_old, u_old, ldf_old, ldu_old
            return
        s0 = self._structs[0]
        s0.f = R.data_ptr()
        s0.ldf = R.stride(0)
        s0.u = W.data_ptr()
        s0.ldu = W.stride(0)
        kernels.mg_vcycle(_lib.F32 if self.dtype == torch.float32 else _lib.F64, self._structs,
                          self.n_levels, self.b, self.nu_pre, self.nu_post, self.coarse_sweeps, self.omega)


class PoissonLatticeMG:
    """Multigrid hierarchy (FP64, one column) for the Poisson stiffness on the
    unknowns of the self-consistent loop's Poisson problem: Dirichlet contacts at
    x = 0 and x = Lx, insulating (natural) boundary at y = 0 and y = Ly, so the
    unknowns are lattice columns 1 .. NX-2 of every row 0 .. NY-1, numbered row
    by row.  Low-order-refined five-point operator with edge weight
    eps * h_perp / h_par (eps averaged over the FE elements on the edge; a
    boundary-row edge has one element and half the weight).  Coarsening is
    vertex-centred: Dirichlet in x (coarse column j at fine 2j + 1), Neumann in y
    (the boundary rows are coarse points: coarse row j at fine 2j), coarse edge
    weight = harmonic mean of the two fine edges it spans.  Used by
    ``qds_cg_mg`` as the CG preconditioner (one V(1,1)-cycle per iteration)."""

    def __init__(self, lattice, eps_elem, min_coarse=3):
        L = lattice
        dev = eps_elem.device
        f64 = dict(dtype=torch.float64, device=dev)
        eps = eps_elem.to(torch.float64)
        NX, NY = L.NX, L.NY
        iy = torch.arange(NY, **f64)[:, None]
        ix = torch.arange(NX - 1, **f64)[None, :]
        # horizontal edges between full columns X and X+1 (X = 0 .. NX-2): (NY, NX-1)
        xm = ((ix + 0.5) * L.hx).expand(NY, NX - 1)
        yy = (iy * L.hy).expand(NY, NX - 1)
        above = eps[L.element_at(xm, yy + 0.25 * L.hy)] * (iy < NY - 1).to(torch.float64)
        below = eps[L.element_at(xm, yy - 0.25 * L.hy)] * (iy > 0).to(torch.float64)
        WE = 0.5 * (above + below) * (L.hy / L.hx)
        # vertical edges between rows Y and Y+1 at the unknown columns X = 1 .. NX-2
        jy = torch.arange(NY - 1, **f64)[:, None]
        jx = torch.arange(1, NX - 1, **f64)[None, :]
        ym = ((jy + 0.5) * L.hy).expand(NY - 1, NX - 2)
        xx = (jx * L.hx).expand(NY - 1, NX - 2)
        WN = 0.5 * (eps[L.element_at(xx + 0.25 * L.hx, ym)] + eps[L.element_at(xx - 0.25 * L.hx, ym)]) * (L.hx / L.hy)
        self.levels_data = []
        while True:
            self.levels_data.append(self._level(WE, WN, dev))
            ny, nx = WE.shape[0], WE.shape[1] - 1
            if not (nx % 2 == 1 and ny % 2 == 1 and (nx - 1) // 2 >= min_coarse and (ny + 1) // 2 >= min_coarse):
                break
            WE, WN = self._coarsen(WE, WN)
        self.n_levels = len(self.levels_data)
        self._structs = (_lib.MGLevel * self.n_levels)()
        for l, lv in enumerate(self.levels_data):
            s = self._structs[l]
            s.nx, s.ny = lv["nx"], lv["ny"]
            s.d, s.we, s.wn = lv["d"].data_ptr(), lv["we"].data_ptr(), lv["wn"].data_ptr()
            s.u, s.f, s.t = lv["u"].data_ptr(), lv["f"].data_ptr(), lv["t"].data_ptr()
            s.ldu = s.ldf = 0
            s.neumann_x, s.neumann_y = 0, 1

    nu, coarse_sweeps, omega = 1, 10, 0.8

    def __call__(self, r, z):
        """z = V(1,1)-cycle(r) for one column (tests; the CG calls the cycle itself)."""
        assert r.dtype == torch.float64 and r.is_contiguous() and z.is_contiguous()
        s0 = self._structs[0]
        f_old, u_old = s0.f, s0.u
        s0.f, s0.u = r.data_ptr(), z.data_ptr()
        try:
            kernels.mg_vcycle(_lib.F64, self._structs, self.n_levels, 1, self.nu, self.nu, self.coarse_sweeps,
                              self.omega)
        finally:
            s0.f, s0.u = f_old, u_old

    @staticmethod
    def _level(WE, WN, dev):
        # WE (ny, nx+1) includes the edges to the Dirichlet columns; WN (ny-1, nx)
        ny, nx = WE.shape[0], WE.shape[1] - 1
        d = WE[:, :-1] + WE[:, 1:]
        d = d.clone()
        d[:-1] += WN
        d[1:] += WN
        we = torch.zeros((ny, nx), dtype=torch.float64, device=dev)
        we[:, :-1] = WE[:, 1:-1]
        wn = torch.zeros((ny, nx), dtype=torch.float64, device=dev)
        wn[:-1] = WN
        n = nx * ny
        return {"nx": nx, "ny": ny, "d": d.contiguous().reshape(-1), "we": we.reshape(-1), "wn": wn.reshape(-1),
                "u": torch.zeros(n, dtype=torch.float64, device=dev),
                "f": torch.zeros(n, dtype=torch.float64, device=dev),
                "t": torch.zeros(n, dtype=torch.float64, device=dev)}

    @staticmethod
    def _coarsen(WE, WN):
        def hmean(a, b):
            s = a + b
            return torch.where(s > 0, 2.0 * a * b / torch.where(s > 0, s, torch.ones_like(s)), torch.zeros_like(s))
        # coarse rows = fine rows 2R; coarse edge E spans fine edges 2E and 2E+1
        WEc = hmean(WE[0::2, 0::2], WE[0::2, 1::2])
        # coarse columns = fine columns 2j+1; coarse vertical edge R spans fine rows 2R, 2R+1
        WNc = hmean(WN[0::2, 1::2], WN[1::2, 1::2])
        return WEc.contiguous(), WNc.contiguous()


class BatchedLatticeMG:
    """LatticeMG for a batch of problems on one lattice that share the kinetic
    coefficient (edge weights) and differ in their potential and shift: every level
    keeps one we / wn and a diagonal per problem, and one qds_mg_vcycle_batched call
    runs the V-cycles of all problems (one launch per pass).  Same construction as
    LatticeMG per problem: level diagonal = edge sums + coarsened (V - sigma) lumped
    mass; the coarse mass term is linear in V, so it is coarsened for all problems at
    once."""

    def __init__(self, lattice, c_elem, v_elems, sigmas, b, dtype=torch.float32, nu_pre=1, nu_post=1,
                 coarse_sweeps=10, omega=0.8, min_coarse=3):
        self.lat = lattice
        self.b = int(b)
        self.dtype = dtype
        self.nu_pre, self.nu_post, self.coarse_sweeps, self.omega = nu_pre, nu_post, coarse_sweeps, omega
        dev = device()
        L = lattice
        c_elem = c_elem.to(dev, torch.float64)
        v_elems = v_elems.to(dev, torch.float64)
        self.batch = B = v_elems.shape[0]
        sig = torch.as_tensor(sigmas, dtype=torch.float64, device=dev).reshape(B)
        f64 = dict(dtype=torch.float64, device=dev)
        ix = torch.arange(L.NX, **f64)
        iy = torch.arange(L.NY, **f64)
        Y, X = torch.meshgrid(iy, ix, indexing="ij")
        xm = (X[:, :-1] + 0.5) * L.hx
        yy = Y[:, :-1] * L.hy
        WE = 0.5 * (c_elem[L.element_at(xm, yy + 0.25 * L.hy)] + c_elem[L.element_at(xm, yy - 0.25 * L.hy)]) \
            * (L.hy / L.hx)
        ym = (Y[:-1, :] + 0.5) * L.hy
        xx = X[:-1, :] * L.hx
        WN = 0.5 * (c_elem[L.element_at(xx + 0.25 * L.hx, ym)] + c_elem[L.element_at(xx - 0.25 * L.hx, ym)]) \
            * (L.hx / L.hy)
        xp = X * L.hx
        yp = Y * L.hy
        vs = sum(v_elems[:, L.element_at(xp + sx * L.hx, yp + sy * L.hy)]
                 for sx in (-0.25, 0.25) for sy in (-0.25, 0.25)) * 0.25          # (B, NY, NX)
        Mterm = (vs - sig[:, None, None]) * (L.hx * L.hy)
        Mass = torch.full((L.NY, L.NX), L.hx * L.hy, **f64)
        self.sigma = sig.clone()
        self.levels_data = []
        NX, NY = L.NX, L.NY
        while True:
            edges = (WE[1:-1, :-1] + WE[1:-1, 1:] + WN[:-1, 1:-1] + WN[1:, 1:-1])
            d64 = edges[None] + Mterm[:, 1:-1, 1:-1]
            we = WE[1:-1, 1:].clone()
            we[:, -1] = 0.0
            wn = WN[1:, 1:-1].clone()
            wn[-1, :] = 0.0
            n = (NX - 2) * (NY - 2)
            self.levels_data.append({
                "nx": NX - 2, "ny": NY - 2, "n": n, "d64": d64.reshape(B, n).contiguous(),
                "d": d64.reshape(B, n).to(dtype).contiguous(),
                "we": we.contiguous().to(dtype).reshape(-1), "wn": wn.contiguous().to(dtype).reshape(-1),
                "mass": Mass[1:-1, 1:-1].reshape(-1).contiguous()})
            nxi, nyi = NX - 2, NY - 2
            if not (nxi % 2 == 1 and nyi % 2 == 1 and (nxi - 1) // 2 >= min_coarse
                    and (nyi - 1) // 2 >= min_coarse):
                break
            WEc, WNc, Massc, NXc, NYc = LatticeMG._coarsen(WE, WN, Mass, NX, NY)
            Mterm = self._restrict_batch(Mterm, NXc, NYc)
            WE, WN, Mass, NX, NY = WEc, WNc, Massc, NXc, NYc
        self.n_levels = len(self.levels_data)
        self._structs = (_lib.MGLevel * self.n_levels)()
        self._strides = (_lib.MGLevelStrides * self.n_levels)()
        for l, lv in enumerate(self.levels_data):
            n = lv["n"]
            for key in ("u", "f", "t"):
                lv[key] = torch.zeros((B, n, self.b), dtype=dtype, device=dev)
            s = self._structs[l]
            s.nx, s.ny = lv["nx"], lv["ny"]
            s.d, s.we, s.wn = lv["d"].data_ptr(), lv["we"].data_ptr(), lv["wn"].data_ptr()
            s.u, s.f, s.t = lv["u"].data_ptr(), lv["f"].data_ptr(), lv["t"].data_ptr()
            s.ldu = s.ldf = 0
            z = self._strides[l]
            z.sd, z.sw, z.su, z.sf, z.st = n, 0, n * self.b, n * self.b, n * self.b

    @staticmethod
    def _restrict_batch(M, NXc, NYc):
        """R m with bilinear weights for (B, NY, NX) (LatticeMG._coarsen per problem)."""
        Mp = torch.nn.functional.pad(M[:, None], (1, 1, 1, 1))[:, 0]
        w = (0.5, 1.0, 0.5)
        Mc = torch.zeros((M.shape[0], NYc, NXc), dtype=M.dtype, device=M.device)
        for a in range(3):
            for c in range(3):
                Mc += w[a] * w[c] * Mp[:, a:a + 2 * NYc - 1:2, c:c + 2 * NXc - 1:2]
        return Mc

    def set_sigma(self, sigmas):
        """Move each problem's shift (diagonals updated in place)."""
        sig = torch.as_tensor(sigmas, dtype=torch.float64, device=self.sigma.device).reshape(self.batch)
        delta = sig - self.sigma
        for lv in self.levels_data:
            lv["d64"] = lv["d64"] - delta[:, None] * lv["mass"][None, :]
            lv["d"].copy_(lv["d64"].to(self.dtype))
        self.sigma = sig.clone()

    @property
    def takes_fp64(self):
        if self.dtype != torch.float32 or self.nu_pre != 1 or self.nu_post != 1 or self.n_levels < 2:
            return False
        lv = self.levels_data[0]
        vec = 4 if self.b % 4 == 0 else 1
        return lv["nx"] * lv["ny"] * (self.b // vec) > 24576

    def __call__(self, R, W):
        """W[i] = V-cycle_i(R[i]) for (batch, n, b) blocks R, W (FP64 blocks on an FP32
        hierarchy when ``takes_fp64``)."""
        assert R.shape == W.shape and R.shape[0] == self.batch and R.is_contiguous() and W.is_contiguous()
        io = _lib.F32 if R.dtype == torch.float32 else _lib.F64
        if R.dtype != self.dtype:
            assert self.dtype == torch.float32 and R.dtype == torch.float64 and self.takes_fp64
        s0, z0 = self._structs[0], self._strides[0]
        saved = (s0.f, s0.u, s0.ldf, s0.ldu, z0.sf, z0.su)
        s0.f, s0.ldf, s0.u, s0.ldu = R.data_ptr(), R.stride(1), W.data_ptr(), W.stride(1)
        z0.sf, z0.su = R.stride(0), W.stride(0)
        try:
            kernels.mg_vcycle_batched(_lib.F32 if self.dtype == torch.float32 else _lib.F64, io, self._structs,
                                      self._strides, self.n_levels, self.batch, self.b, self.nu_pre, self.nu_post,
                                      self.coarse_sweeps, self.omega)
        finally:
            s0.f, s0.u, s0.ldf, s0.ldu, z0.sf, z0.su = saved
