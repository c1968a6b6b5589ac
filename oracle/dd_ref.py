"""FP64 oracle of the full Poisson--drift--diffusion solver -- TEST INFRASTRUCTURE ONLY.

Restates FullPoissonDriftDiffusionSolver (backend/src/full_poisson_dd_solver.cpp)
on a P1 triangle mesh, in numpy / scipy, function by function:

  material lookup ................. :185-211 (GaAs defaults, first matching region)
  initial carriers ................ :216-260
  charge density .................. :267-288  rho = q (p - n + N)
  Poisson step .................... :296-519  K = eps grad.grad area (eps at the centroid, F/cm),
                                               f_i -= sum_q rho_q lambda_i(q) (1/6) area,
                                               Dirichlet V_p (x < 1e-6) / V_n (x > Lx - 1e-6)
                                               by row replacement, NaN -> linear ramp
  node-averaged field ............. :524-583  E = -grad phi, plain average over the node's elements
  field at a point ................ :592-651  first element with all barycentrics >= 0, else 0
  Boltzmann / Fermi-Dirac ......... :656-718  clamp max(., 1e5) with std::max's NaN behaviour
  mobilities, G-R ................. :61-82, :726-770 (constant mobility, SRH with ni = 1.5e6)
  contact conditions .............. :778-839  ohmic, N_a = N_d = 1e17
  continuity step ................. :847-1063 D grad.grad area + (+/-) mu E.grad N_j area / 3,
                                               load G_c area / 3, Dirichlet rows, LDLT solve
  currents ........................ :1068-1159 J = q mu_c n_c grad phi_n, averaged per node
  damping ......................... :1168-1173
  Gummel loop ..................... :1296-1427

Linear systems.  The reference factorises every system with
Eigen::SimplicialLDLT<SparseMatrix<double>>, whose default UpLo = Lower reads only
the lower triangle: the matrix it factors is S = tril(A) + tril(A, -1)^T of the
row-replaced A, which is not A for the nonsymmetric continuity matrices nor for
the row-replaced Poisson matrix.  ``ldlt_lower=True`` solves S (what the reference
computes); ``ldlt_lower=False`` (default) solves A itself, i.e. the Dirichlet
problems the code writes down.  Both use scipy's spsolve.

Pinning (tests/test_dd_oracle.py): the Poisson step against the nodally exact
quadratic for uniform charge (fixes the sign, the 1/6-weight factor and epsilon_0);
the field and the currents of linear potentials (exact); the statistics round trip
of the initial carriers (Boltzmann with phi = 0 returns the mass-action carriers);
the continuity operator through K 1 = 0, the divergence-theorem identity
1^T K_drift x = +/- mu E_x |Omega| (fixes the drift sign and which index carries the
gradient), the E = 0 operator against fem_ref's P1 stiffness, the load sum
G |Omega|; the continuity solve against the nodally exact parabola
phi_n = linear + G/(2D) x (L - x) for E = 0 and uniform G; S against np.tril of the
dense matrix.  The reading of SimplicialLDLT as lower-triangle-only is Eigen's
documented default and is not executed here (no Eigen in this image).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

KT = 0.0259
Q = 1.602e-19
EPS0_F_PER_CM = 8.85418782e-14
N_MIN = 1e5
BC_TOL = 1e-6
DAMPING = 0.3
CONTACT_DOPING = 1e17
SRH_NI = 1.5e6
SRH_TAU_N = 1e-9
SRH_TAU_P = 1e-9

GAAS = dict(N_c=4.7e17, N_v=7.0e18, E_g=1.424, mu_n=8500.0, mu_p=400.0, epsilon_r=12.9, m_e=0.067, m_h=0.45)


def material_index(x, y, regions):
    """Index of the material at each point: the first region that holds it, else 0
    (:185-211); with no regions every point uses the GaAs defaults (index 0)."""
    idx = np.zeros(len(x), dtype=np.int64)
    if not regions:
        return idx
    for k in range(len(x)):
        for r, f in enumerate(regions):
            if f(float(x[k]), float(y[k])):
                idx[k] = r
                break
    return idx


def std_max(a, b):
    """std::max(a, b) = (a < b) ? b : a, elementwise (NaN in a stays NaN)."""
    return np.where(a < b, b, a)


def geometry(x, y, conn):
    """Areas and shape-function gradients grad_N[e, i, (x, y)] = J^-T ref_grad_i (:322-345)."""
    x1, x2, x3 = (x[conn[:, k]] for k in range(3))
    y1, y2, y3 = (y[conn[:, k]] for k in range(3))
    area = 0.5 * np.abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1))
    J = np.empty((len(conn), 2, 2))
    J[:, 0, 0], J[:, 0, 1] = x2 - x1, x3 - x1
    J[:, 1, 0], J[:, 1, 1] = y2 - y1, y3 - y1
    Jit = np.transpose(np.linalg.inv(J), (0, 2, 1))
    ref = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
    grad = np.einsum("eab,ib->eia", Jit, ref)
    return area, grad


def node_gradient_average(x, y, conn, u, weight=None):
    """Per node: mean over the node's elements of weight_e * grad u_e, the gradient
    from dN/dx = (y2 - y3) / 2A ... (:544-560, :1095-1128).  Nodes in no element get 0."""
    x1, x2, x3 = (x[conn[:, k]] for k in range(3))
    y1, y2, y3 = (y[conn[:, k]] for k in range(3))
    A2 = 2.0 * (0.5 * np.abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)))
    dx = np.stack([(y2 - y3), (y3 - y1), (y1 - y2)], 1) / A2[:, None]
    dy = np.stack([(x3 - x2), (x1 - x3), (x2 - x1)], 1) / A2[:, None]
    ue = u[conn]
    g = np.stack([(ue * dx).sum(1), (ue * dy).sum(1)], 1)
    if weight is not None:
        g = g * weight[:, None]
    n = len(x)
    acc = np.zeros((n, 2))
    cnt = np.zeros(n)
    for k in range(3):
        np.add.at(acc, conn[:, k], g)
        np.add.at(cnt, conn[:, k], 1.0)
    out = np.zeros((n, 2))
    nz = cnt > 0
    out[nz] = acc[nz] / cnt[nz, None]
    return out


def ldlt_lower_matrix(A):
    """The symmetric matrix Eigen's SimplicialLDLT<., Lower> factors for A:
    tril(A) + tril(A, -1)^T."""
    L = sp.tril(A, format="csr")
    return (L + sp.tril(A, k=-1, format="csr").T).tocsr()


def _solve(A, f, ldlt_lower):
    M = ldlt_lower_matrix(A) if ldlt_lower else A.tocsc()
    return spla.spsolve(M.tocsc(), f)


def _row_replace(A, rows, f, values):
    A = A.tolil()
    for i in rows:
        A.rows[i] = [i]
        A.data[i] = [1.0]
    f = f.copy()
    f[rows] = values
    return A.tocsr(), f


class DDRef:
    """State and steps of FullPoissonDriftDiffusionSolver on one P1 mesh."""

    def __init__(self, x, y, conn, Lx, doping, materials=None, regions=None, fermi_dirac=False,
                 mobility=None, g_r=None, ldlt_lower=False):
        self.x, self.y = np.asarray(x, float), np.asarray(y, float)
        self.conn = np.asarray(conn, np.int64)
        self.Lx = float(Lx)
        self.n_nodes = len(self.x)
        self.doping = np.asarray(doping, float)
        mats = materials if materials else [GAAS]
        self.mats = [m if isinstance(m, dict) else {k: getattr(m, k) for k in GAAS} for m in mats]
        self.mat_objs = list(mats)
        idx = material_index(self.x, self.y, regions if materials else None)
        # centroid (x1 + x2 + x3) / 3 as in :344-345
        xc = (self.x[self.conn[:, 0]] + self.x[self.conn[:, 1]] + self.x[self.conn[:, 2]]) / 3.0
        yc = (self.y[self.conn[:, 0]] + self.y[self.conn[:, 1]] + self.y[self.conn[:, 2]]) / 3.0
        eidx = material_index(xc, yc, regions if materials else None)
        self.node_mat = idx
        self.par = {k: np.array([self.mats[i][k] for i in idx]) for k in GAAS}
        self.eps_e = np.array([self.mats[i]["epsilon_r"] for i in eidx]) * EPS0_F_PER_CM
        self.fermi_dirac = fermi_dirac
        self.mobility = mobility        # None: constant mat.mu_n / mat.mu_p
        self.g_r = g_r                  # None: SRH of :72-82
        self.ldlt_lower = ldlt_lower
        self.area, self.grad = geometry(self.x, self.y, self.conn)
        self.left = np.nonzero(self.x < BC_TOL)[0]
        self.right = np.nonzero((self.x > self.Lx - BC_TOL) & ~(self.x < BC_TOL))[0]
        n = self.n_nodes
        self.phi = np.zeros(n)
        self.E = np.zeros((n, 2))
        self.Jn = np.zeros((n, 2))
        self.Jp = np.zeros((n, 2))
        self.initialize()
        self.poisson_solves = 0

    # -- :216-260 ------------------------------------------------------------------
    def ni(self):
        P = self.par
        return np.sqrt(P["N_c"] * P["N_v"]) * np.exp(-P["E_g"] / (2.0 * KT))

    def initialize(self):
        P, N = self.par, self.doping
        ni = self.ni()
        with np.errstate(divide="ignore", invalid="ignore"):
            n = np.where(N > 0, N, np.where(N < 0, ni * ni / -N, ni))
            p = np.where(N > 0, ni * ni / N, np.where(N < 0, -N, ni))
        self.n = std_max(n, N_MIN)
        self.p = std_max(p, N_MIN)
        self.phi_n = KT * np.log(self.n / P["N_c"])
        self.phi_p = -KT * np.log(self.p / P["N_v"]) - P["E_g"]

    # -- :267-288 ------------------------------------------------------------------
    def charge_density(self):
        return Q * (self.p - self.n + self.doping)

    # -- :296-519 ------------------------------------------------------------------
    def poisson_matrix(self):
        c = self.conn
        Ke = self.eps_e[:, None, None] * np.einsum("eia,eja->eij", self.grad, self.grad) * self.area[:, None, None]
        rows = np.repeat(c, 3, axis=1).ravel()
        cols = np.tile(c, (1, 3)).ravel()
        return sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(self.n_nodes,) * 2).tocsr()

    def poisson_load(self, rho):
        lam = np.array([[1 / 6, 1 / 6, 2 / 3], [1 / 6, 2 / 3, 1 / 6], [2 / 3, 1 / 6, 1 / 6]])
        f = np.zeros(self.n_nodes)
        re = rho[self.conn]
        for q in range(3):
            rq = re @ lam[q]
            for i in range(3):
                np.add.at(f, self.conn[:, i], -rq * lam[q, i] * (1.0 / 6.0) * self.area)
        return f

    def poisson_step(self, V_p, V_n):
        K = self.poisson_matrix()
        f = self.poisson_load(self.charge_density())
        self.phi[self.left] = V_p
        self.phi[self.right] = V_n
        bc = np.concatenate([self.left, self.right])
        vals = np.concatenate([np.full(len(self.left), float(V_p)), np.full(len(self.right), float(V_n))])
        A, f = _row_replace(K, bc, f, vals)
        phi = _solve(A, f, self.ldlt_lower)
        if not np.all(np.isfinite(phi)):
            phi = V_p + (V_n - V_p) * self.x / self.Lx
        self.phi = phi
        self.poisson_solves += 1
        self.electric_field()

    # -- :524-583 ------------------------------------------------------------------
    def electric_field(self):
        self.E = -node_gradient_average(self.x, self.y, self.conn, self.phi)

    # -- :592-651 ------------------------------------------------------------------
    def field_at(self, px, py):
        x1, x2, x3 = (self.x[self.conn[:, k]] for k in range(3))
        y1, y2, y3 = (self.y[self.conn[:, k]] for k in range(3))
        A2 = 2.0 * (0.5 * np.abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)))
        l1 = ((y2 - y3) * (px - x3) + (x3 - x2) * (py - y3)) / A2
        l2 = ((y3 - y1) * (px - x3) + (x1 - x3) * (py - y3)) / A2
        l3 = 1.0 - l1 - l2
        hit = np.nonzero((l1 >= 0) & (l2 >= 0) & (l3 >= 0))[0]
        if hit.size == 0:
            return np.zeros(2)
        e = hit[0]
        sub = self.conn[e:e + 1]
        g = -node_gradient_average(self.x, self.y, sub, self.phi)[sub[0, 0]]
        return g

    # -- :656-718 ------------------------------------------------------------------
    def carriers(self):
        P = self.par
        eta_n = (self.phi_n - self.phi) / KT
        eta_p = (self.phi - self.phi_p - P["E_g"]) / KT
        with np.errstate(over="ignore", invalid="ignore"):
            if self.fermi_dirac:
                n = P["N_c"] * fermi_dirac_half(eta_n)
                p = P["N_v"] * fermi_dirac_half(eta_p)
            else:
                n = P["N_c"] * np.exp(eta_n)
                p = P["N_v"] * np.exp(eta_p)
        self.n = std_max(n, N_MIN)
        self.p = std_max(p, N_MIN)

    # -- :61-82, :726-770 -------------------------------------------------------------
    def mobilities(self):
        if self.mobility is None:
            return self.par["mu_n"].copy(), self.par["mu_p"].copy()
        Em = np.sqrt((self.E ** 2).sum(1))
        mu_n = np.array([self.mobility[0](self.x[i], self.y[i], Em[i], self.mat_objs[self.node_mat[i]])
                         for i in range(self.n_nodes)], float)
        mu_p = np.array([self.mobility[1](self.x[i], self.y[i], Em[i], self.mat_objs[self.node_mat[i]])
                         for i in range(self.n_nodes)], float)
        return mu_n, mu_p

    def generation_recombination(self):
        if self.g_r is None:
            return srh(self.n, self.p)
        return np.array([self.g_r(self.x[i], self.y[i], self.n[i], self.p[i], self.mat_objs[self.node_mat[i]])
                         for i in range(self.n_nodes)], float)

    # -- :778-839 ------------------------------------------------------------------
    def apply_boundary_conditions(self, V_p, V_n):
        P = self.par
        ni = self.ni()
        L, R = self.left, self.right
        self.phi[L] = V_p
        self.phi_n[L] = V_p + KT * np.log(ni[L] * ni[L] / CONTACT_DOPING / P["N_c"][L])
        self.phi_p[L] = V_p - P["E_g"][L]
        self.n[L] = ni[L] * ni[L] / CONTACT_DOPING
        self.p[L] = CONTACT_DOPING
        self.phi[R] = V_n
        self.phi_n[R] = V_n
        self.phi_p[R] = V_n - P["E_g"][R] - KT * np.log(CONTACT_DOPING / P["N_v"][R])
        self.n[R] = CONTACT_DOPING
        self.p[R] = ni[R] * ni[R] / CONTACT_DOPING

    # -- :847-1063 -----------------------------------------------------------------
    def continuity_matrices(self, mu_n, mu_p, G):
        c = self.conn
        third = 1.0 / 3.0
        Ec = sum(third * self.E[c[:, k]] for k in range(3))
        mun_c = sum(third * mu_n[c[:, k]] for k in range(3))
        mup_c = sum(third * mu_p[c[:, k]] for k in range(3))
        G_c = sum(third * G[c[:, k]] for k in range(3))
        gg = np.einsum("eia,eja->eij", self.grad, self.grad) * self.area[:, None, None]
        eg = np.einsum("ea,eja->ej", Ec, self.grad)            # E_c . grad N_j
        Kn_e = (mun_c * KT)[:, None, None] * gg + (mun_c * self.area / 3.0)[:, None, None] * eg[:, None, :]
        Kp_e = (mup_c * KT)[:, None, None] * gg - (mup_c * self.area / 3.0)[:, None, None] * eg[:, None, :]
        rows = np.repeat(c, 3, axis=1).ravel()
        cols = np.tile(c, (1, 3)).ravel()
        shape = (self.n_nodes,) * 2
        Kn = sp.coo_matrix((Kn_e.ravel(), (rows, cols)), shape=shape).tocsr()
        Kp = sp.coo_matrix((Kp_e.ravel(), (rows, cols)), shape=shape).tocsr()
        f = np.zeros(self.n_nodes)
        for i in range(3):
            np.add.at(f, c[:, i], G_c * self.area / 3.0)
        return Kn, Kp, f

    def drift_diffusion_step(self, V_p, V_n):
        mu_n, mu_p = self.mobilities()
        G = self.generation_recombination()
        Kn, Kp, f = self.continuity_matrices(mu_n, mu_p, G)
        self.apply_boundary_conditions(V_p, V_n)
        bc = np.concatenate([self.left, self.right])
        An, fn = _row_replace(Kn, bc, f, self.phi_n[bc])
        Ap, fp = _row_replace(Kp, bc, f, self.phi_p[bc])
        self.phi_n = _solve(An, fn, self.ldlt_lower)
        self.phi_p = _solve(Ap, fp, self.ldlt_lower)
        self.carriers()
        self.currents()

    # -- :1068-1159 ----------------------------------------------------------------
    def currents(self):
        mu_n, mu_p = self.mobilities()
        c = self.conn
        third = 1.0 / 3.0
        n_c = sum(third * self.n[c[:, k]] for k in range(3))
        p_c = sum(third * self.p[c[:, k]] for k in range(3))
        mun_c = sum(third * mu_n[c[:, k]] for k in range(3))
        mup_c = sum(third * mu_p[c[:, k]] for k in range(3))
        self.Jn = node_gradient_average(self.x, self.y, c, self.phi_n, Q * mun_c * n_c)
        self.Jp = node_gradient_average(self.x, self.y, c, self.phi_p, Q * mup_c * p_c)

    # -- :1296-1427 ----------------------------------------------------------------
    def solve(self, V_p, V_n, tolerance=1e-6, max_iter=100, min_iter=0):
        """The reference loop; ``min_iter`` forces that many iterations before the
        stopping test (0 = the reference)."""
        self.initialize()
        self.apply_boundary_conditions(V_p, V_n)
        self.phi = V_p + (V_n - V_p) * self.x / self.Lx
        self.electric_field()
        self.carriers()
        self.electric_field()
        self.currents()
        error, it = 1.0, 0
        while (error > tolerance or it < min_iter) and it < max_iter:
            phi_old = self.phi.copy()
            if it > 0:
                self.poisson_step(V_p, V_n)
                if it > 1:
                    self.phi = phi_old + DAMPING * (self.phi - phi_old)
            self.electric_field()
            self.drift_diffusion_step(V_p, V_n)
            bad = ~np.isfinite(self.phi)
            if bad.any():
                self.phi[bad] = phi_old[bad]
                error = 1.0
            else:
                pn = np.linalg.norm(self.phi)
                dn = np.linalg.norm(self.phi - phi_old)
                error = dn / pn if pn > 1e-10 else dn
            if not np.isfinite(error):
                error = 1.0
            it += 1
        self.iterations = it
        self.error = error
        self.electric_field()
        self.currents()
        return it, error


def fermi_dirac_half(eta):
    """F_1/2 approximation of :690-697."""
    eta = np.asarray(eta, float)
    with np.errstate(over="ignore", invalid="ignore"):
        ex = np.exp(np.minimum(eta, 0.0))
        pos = np.maximum(eta, 0.0) ** 1.5
        return np.where(eta <= 0, ex / (1.0 + 0.27 * ex), pos / (1.5 + 0.27 * pos))


def srh(n, p):
    """Default generation-recombination rate (:72-82)."""
    ni = SRH_NI
    return (n * p - ni * ni) / (SRH_TAU_P * (n + ni) + SRH_TAU_N * (p + ni))
