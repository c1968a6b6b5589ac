"""Poisson--drift--diffusion on the GPU.

Drop-in for FullPoissonDriftDiffusionSolver (backend/src/full_poisson_dd_solver.cpp,
binding backend/src/bindings.cpp:1611-1650): ``FullPoissonDriftDiffusionSolver(mesh,
epsilon_r, doping_profile)``, ``solve(V_p, V_n, tolerance=1e-6, max_iter=100)``,
``get_potential``, ``get_electron_concentration``, ``get_hole_concentration``,
``get_electric_field(x, y)``, ``get_electric_field_all``,
``get_electron_current_density``, ``get_hole_current_density``,
``set_heterojunction``, ``set_generation_recombination_model``,
``set_mobility_models``, ``set_carrier_statistics_model``,
``enable_quantum_corrections``, ``enable_adaptive_mesh_refinement``.

Units and constants are the reference's: coordinates as the mesh holds them (the
reference's examples use nm), concentrations in cm^-3, kT = 0.0259 eV,
q = 1.602e-19 C, eps0 = 8.85418782e-14 F/cm, permittivity from the material at
the element centroid (the epsilon_r callback is kept but, as in the reference, not
used by the Poisson step), GaAs defaults (:185-211).  Contacts are the nodes with
x < 1e-6 (V_p) and x > Lx - 1e-6 (V_n) in mesh units.

The Gummel loop is the reference's (:1296-1427): iteration 0 solves only the
continuity equations on the linear potential ramp; later iterations solve Poisson
first and damp the potential by 0.3 from iteration 2 on; the loop stops when
|phi - phi_old| / |phi| <= tolerance.  Since iteration 0 leaves phi unchanged, the
reference stops after that one iteration whenever the potential stays finite;
``min_iter`` (keyword, default 0 = the reference) forces more iterations.

Each step runs on the device (dd.hip): node kernels for the carrier statistics,
contacts and charge; node averages for the field and the currents; row kernels that
write the continuity and Poisson matrices straight into the P1 pattern with the
Dirichlet rows replaced.  Linear systems:

* default (``ldlt_lower=False``): the Dirichlet problems as the code writes them,
  restricted to the free (non-contact) nodes with the contact values lifted into
  the load.  The Poisson step is solved by CG and the continuity systems by
  BiCGStab, both preconditioned by a multigrid V-cycle on structured meshes whose
  contacts are the first and last lattice columns (PoissonProblem, qds_cg_mg,
  qds_bicgstab_mg; the continuity hierarchies discretise the diffusion part
  D grad.grad).  Other meshes: Jacobi-CG, and Jacobi BiCGStab on the row-replaced
  matrices (qds_bicgstab).
* ``ldlt_lower=True``: the systems the reference's SimplicialLDLT (UpLo = Lower)
  factors, i.e. the lower triangle of each row-replaced matrix mirrored, solved by
  BiCGStab.  This reproduces the reference's numbers; the default solves the
  equations the code states.

Iterative solves stop at a relative residual ``rtol`` (1e-12); one that stops short
is counted in ``last_info`` and warned about.  Custom mobility and
generation-recombination callbacks (``set_mobility_models``,
``set_generation_recombination_model``) are Python callables evaluated on the host
per node, as in the reference; the default models run in the kernels.
"""
import ctypes
import math
import warnings
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, kernels
from .device import device, ptr, stream_ptr, workspace
from .poisson import PoissonProblem

KT = 0.0259
Q = 1.602e-19
EPS0_F_PER_CM = 8.85418782e-14
N_MIN = 1e5
BC_TOL = 1e-6
DAMPING = 0.3
CONTACT_DOPING = 1e17
NM = 1e-9


@dataclass
class Material:
    """Material parameters (Materials::Material); defaults are the GaAs values of
    get_material_at (:185-200)."""
    N_c: float = 4.7e17
    N_v: float = 7.0e18
    E_g: float = 1.424
    mu_n: float = 8500.0
    mu_p: float = 400.0
    epsilon_r: float = 12.9
    m_e: float = 0.067
    m_h: float = 0.45


def _call_points(func, x, y, *extra):
    """Evaluate func over arrays (one vectorised call if it returns the right shape,
    else a per-point loop)."""
    n = len(x)
    try:
        v = np.asarray(func(x, y, *extra), dtype=np.float64)
        if v.shape == (n,):
            return v
        if v.shape == ():
            return np.full(n, float(v))
    except Exception:
        pass
    cols = [np.broadcast_to(np.asarray(a), (n,)) if not hasattr(a, "N_c") else None for a in extra]
    out = np.empty(n)
    for k in range(n):
        args = [c[k] if c is not None else a for c, a in zip(cols, extra)]
        out[k] = float(func(float(x[k]), float(y[k]), *args))
    return out


class FullPoissonDriftDiffusionSolver:
    """GPU Poisson--drift--diffusion on a P1 mesh (see the module docstring)."""

    def __init__(self, mesh, epsilon_r, doping_profile, *, ldlt_lower=False, rtol=1e-12, max_krylov=20000):
        if int(getattr(mesh, "element_order", 1)) != 1:
            raise ValueError("FullPoissonDriftDiffusionSolver works on linear (3-node) elements")
        self.mesh = mesh
        self.epsilon_r = epsilon_r
        self.doping_profile = doping_profile
        self.ldlt_lower = bool(ldlt_lower)
        self.rtol = float(rtol)
        self.max_krylov = int(max_krylov)
        self.dev = dev = device()
        x, y = mesh.get_nodes()
        self.x_host = np.ascontiguousarray(x, dtype=np.float64)
        self.y_host = np.ascontiguousarray(y, dtype=np.float64)
        conn = np.ascontiguousarray(mesh.get_elements(), dtype=np.int32)
        if conn.ndim != 2 or conn.shape[1] != 3:
            raise ValueError("FullPoissonDriftDiffusionSolver needs (E, 3) connectivity")
        self.conn_host = conn
        self.n = n = int(mesh.num_nodes)
        self.n_elem = int(conn.shape[0])
        self.Lx = float(mesh.Lx)
        # node -> elements in ascending element id
        flat = conn.ravel().astype(np.int64)
        order = np.argsort(flat, kind="stable")
        ne_ptr = np.zeros(n + 1, np.int64)
        np.add.at(ne_ptr, flat + 1, 1)
        ne_ptr = np.cumsum(ne_ptr).astype(np.int32)
        ne_elem = (order // 3).astype(np.int32)
        pat = kernels.pattern_build(conn, n)
        self.indptr_host, self.indices_host = pat["indptr"], pat["indices"]
        self.nnz = int(pat["nnz"])
        contact = np.zeros(n, np.int8)
        contact[self.x_host < BC_TOL] = 1
        contact[(self.x_host > self.Lx - BC_TOL) & (contact == 0)] = 2
        self.contact_host = contact
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self._x, self._y = t(self.x_host), t(self.y_host)
        self._conn = t(conn)
        self._ne_ptr, self._ne_elem = t(ne_ptr), t(ne_elem)
        self._indptr, self._indices = t(pat["indptr"]), t(pat["indices"])
        self._contact = t(contact)
        doping = _call_points(doping_profile, self.x_host, self.y_host)
        self.doping_host = doping
        self._doping = t(doping)
        z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
        self._phi, self._phi_n, self._phi_p, self._n, self._p = z(), z(), z(), z(), z()
        self._phi_old = z()
        self._rho = z()
        self._fn, self._fp = z(), z()
        self._E = torch.zeros((n, 2), dtype=torch.float64, device=dev)
        self._Jn = torch.zeros((n, 2), dtype=torch.float64, device=dev)
        self._Jp = torch.zeros((n, 2), dtype=torch.float64, device=dev)
        self._Kn = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        self._Kp = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        self._K = None
        self._sums = torch.zeros(4, dtype=torch.float64, device=dev)
        self.use_fermi_dirac_statistics = False
        self.use_quantum_corrections = False
        self.use_adaptive_mesh_refinement = False
        self.adaptive_mesh_refinement_threshold = 0.1
        self.adaptive_mesh_refinement_max_level = 3
        self.damping_factor = DAMPING
        self.mu_n_model = None          # None: constant mat.mu_n (:61-63)
        self.mu_p_model = None
        self.g_r_model = None           # None: SRH (:72-82)
        self.materials = [Material()]
        self.regions = []
        self.has_heterojunction = False
        self.iterations = 0
        self.error = float("nan")
        self.last_info = {}
        self._poisson = None
        self._cont_mg = None
        self._set_materials()
        self.initialize_carrier_concentrations()

    # -- materials ----------------------------------------------------------------------
    def _material_index(self, px, py):
        idx = np.zeros(len(px), np.int32)
        if not self.has_heterojunction:
            return idx
        for k in range(len(px)):
            for r, f in enumerate(self.regions):
                if f(float(px[k]), float(py[k])):
                    idx[k] = r
                    break
        return idx

    def _set_materials(self):
        mats = self.materials if self.has_heterojunction else [Material()]
        self._mats = mats
        table = np.zeros((len(mats), 8))
        for i, m in enumerate(mats):
            ni = math.sqrt(m.N_c * m.N_v) * math.exp(-m.E_g / (2.0 * KT))
            table[i] = [m.N_c, m.N_v, m.E_g, m.mu_n, m.mu_p, m.epsilon_r, ni, 0.0]
        c = self.conn_host
        xc = (self.x_host[c[:, 0]] + self.x_host[c[:, 1]] + self.x_host[c[:, 2]]) / 3.0
        yc = (self.y_host[c[:, 0]] + self.y_host[c[:, 1]] + self.y_host[c[:, 2]]) / 3.0
        self.node_mat = self._material_index(self.x_host, self.y_host)
        self.elem_mat = self._material_index(xc, yc)
        dev = self.dev
        self._mat = torch.from_numpy(table.ravel().copy()).to(dev)
        self._mat_idx = torch.from_numpy(self.node_mat).to(dev)
        self._emat_idx = torch.from_numpy(self.elem_mat).to(dev)
        self._mu_n_const = torch.from_numpy(table[self.node_mat, 3].copy()).to(dev)
        self._mu_p_const = torch.from_numpy(table[self.node_mat, 4].copy()).to(dev)
        self._eps_elem = table[self.elem_mat, 5] * EPS0_F_PER_CM
        self._m = _lib.DDMesh(self.n, self.n_elem, ptr(self._conn), ptr(self._x), ptr(self._y), ptr(self._ne_ptr),
                              ptr(self._ne_elem), ptr(self._indptr), ptr(self._indices), ptr(self._mat_idx),
                              ptr(self._emat_idx), ptr(self._mat), ptr(self._contact))
        self._poisson = None

    def _poisson_problem(self):
        """Free-node Poisson operator (default mode), built on first use; its free-node
        numbering, restricted pattern and slot map also serve the continuity systems."""
        if self._poisson is None:
            left = np.nonzero(self.contact_host == 1)[0]
            right = np.nonzero(self.contact_host == 2)[0]
            self._poisson = PoissonProblem(self.mesh, None, sign="reference", eps_elem=self._eps_elem,
                                           contacts=(left, right))
            self._cont_mg = None
            if self._poisson.mg is not None:
                # preconditioners of the continuity systems: multigrid hierarchies of
                # their diffusion parts D_e grad.grad (D_e = kT x the element mean of
                # the materials' mobilities); the drift part, and any field-dependent
                # mobility model, is left to the Krylov solver
                from .lattice import Lattice, PoissonLatticeMG
                lat = Lattice(self.mesh)
                c = self.conn_host
                mats = np.array([[m.mu_n, m.mu_p] for m in self._mats])
                mu = mats[self.node_mat]
                d = KT * (mu[c[:, 0]] + mu[c[:, 1]] + mu[c[:, 2]]) / 3.0
                self._cont_mg = tuple(PoissonLatticeMG(lat, torch.from_numpy(d[:, k].copy()).to(self.dev))
                                      for k in range(2))
            self._mask = torch.from_numpy((self.contact_host != 0).astype(np.float64)).to(self.dev)
        return self._poisson

    # -- setters (:97-176) ------------------------------------------------------------------
    def set_heterojunction(self, materials, regions):
        if len(materials) != len(regions):
            raise ValueError("Number of materials must match number of regions")
        self.materials = list(materials)
        self.regions = list(regions)
        self.has_heterojunction = True
        self._set_materials()
        self.initialize_carrier_concentrations()

    def set_generation_recombination_model(self, g_r):
        self.g_r_model = g_r

    def set_mobility_models(self, mu_n, mu_p):
        self.mu_n_model, self.mu_p_model = mu_n, mu_p

    def set_carrier_statistics_model(self, use_fermi_dirac):
        self.use_fermi_dirac_statistics = bool(use_fermi_dirac)
        self.initialize_carrier_concentrations()

    def enable_quantum_corrections(self, enable):
        # stored only, as in the reference (:159-161)
        self.use_quantum_corrections = bool(enable)

    def enable_adaptive_mesh_refinement(self, enable, refinement_threshold=0.1, max_refinement_level=3):
        # stored only, as in the reference (:170-176)
        self.use_adaptive_mesh_refinement = bool(enable)
        self.adaptive_mesh_refinement_threshold = refinement_threshold
        self.adaptive_mesh_refinement_max_level = max_refinement_level

    # -- device steps ---------------------------------------------------------------------
    def initialize_carrier_concentrations(self):
        _lib.check(_lib.lib.qds_dd_initial(ctypes.byref(self._m), KT, N_MIN, ptr(self._doping), ptr(self._n),
                                           ptr(self._p), ptr(self._phi_n), ptr(self._phi_p), stream_ptr()),
                   "qds_dd_initial")

    def _contacts(self, V_p, V_n, loads=False):
        _lib.check(_lib.lib.qds_dd_contacts(ctypes.byref(self._m), float(V_p), float(V_n), KT, CONTACT_DOPING,
                                            ptr(self._phi), ptr(self._phi_n), ptr(self._phi_p), ptr(self._n),
                                            ptr(self._p), ptr(self._fn) if loads else None,
                                            ptr(self._fp) if loads else None, stream_ptr()), "qds_dd_contacts")

    def _carriers(self):
        _lib.check(_lib.lib.qds_dd_carriers(ctypes.byref(self._m), int(self.use_fermi_dirac_statistics), KT, N_MIN,
                                            ptr(self._phi), ptr(self._phi_n), ptr(self._phi_p), ptr(self._n),
                                            ptr(self._p), stream_ptr()), "qds_dd_carriers")

    def compute_electric_field(self):
        _lib.check(_lib.lib.qds_dd_field(ctypes.byref(self._m), ptr(self._phi), ptr(self._E), stream_ptr()),
                   "qds_dd_field")

    def _mobilities(self):
        if self.mu_n_model is None and self.mu_p_model is None:
            return self._mu_n_const, self._mu_p_const
        Em = torch.linalg.vector_norm(self._E, dim=1).cpu().numpy()
        out = []
        for model, const in ((self.mu_n_model, self._mu_n_const), (self.mu_p_model, self._mu_p_const)):
            if model is None:
                out.append(const)
                continue
            vals = np.empty(self.n)
            for mi, mat in enumerate(self._mats):
                sel = np.nonzero(self.node_mat == mi)[0]
                if sel.size:
                    vals[sel] = _call_points(model, self.x_host[sel], self.y_host[sel], Em[sel], mat)
            out.append(torch.from_numpy(vals).to(self.dev))
        return out[0], out[1]

    def _generation_recombination(self):
        if self.g_r_model is None:
            return None
        nh, ph = self._n.cpu().numpy(), self._p.cpu().numpy()
        vals = np.empty(self.n)
        for mi, mat in enumerate(self._mats):
            sel = np.nonzero(self.node_mat == mi)[0]
            if sel.size:
                vals[sel] = _call_points(self.g_r_model, self.x_host[sel], self.y_host[sel], nh[sel], ph[sel], mat)
        return torch.from_numpy(vals).to(self.dev)

    def compute_current_densities(self, mu=None):
        mu_n, mu_p = mu if mu is not None else self._mobilities()
        _lib.check(_lib.lib.qds_dd_currents(ctypes.byref(self._m), Q, ptr(self._phi_n), ptr(self._phi_p), ptr(self._n),
                                            ptr(self._p), ptr(mu_n), ptr(mu_p), ptr(self._Jn), ptr(self._Jp),
                                            stream_ptr()), "qds_dd_currents")

    def _krylov(self, vals, b, x, what):
        it, rr, rs = kernels.bicgstab(self._indptr, self._indices, vals, self.n, self.nnz, b, x, rtol=self.rtol,
                                      max_iter=self.max_krylov)
        return self._record(what, it, rr, rs)

    def _record(self, what, it, rr, rs):
        info = {"iterations": it, "relres": rr, "restarts": rs}
        self.last_info[what] = info
        if not rr <= self.rtol:
            self.last_info["unconverged"] = self.last_info.get("unconverged", 0) + 1
            warnings.warn(f"{what}: BiCGStab stopped after {it} iterations at relative residual "
                          f"{rr:.3e} (rtol {self.rtol:.1e})", RuntimeWarning, stacklevel=3)
        return info

    def poisson_system(self, V_p, V_n, ldlt_lower=None):
        """(K, f) of the Poisson step on the mesh pattern (CSR values, device), with
        the Dirichlet rows replaced; ldlt_lower (default: the solver's mode) mirrors
        the lower triangle as the reference's factorisation reads it."""
        ldlt = self.ldlt_lower if ldlt_lower is None else bool(ldlt_lower)
        if self._K is None:
            self._K = torch.empty(self.nnz, dtype=torch.float64, device=self.dev)
        _lib.check(_lib.lib.qds_dd_poisson_system(ctypes.byref(self._m), int(ldlt), EPS0_F_PER_CM, Q, ptr(self._n),
                                                  ptr(self._p), ptr(self._doping), float(V_p), float(V_n),
                                                  ptr(self._K), ptr(self._rho), stream_ptr()),
                   "qds_dd_poisson_system")
        return self._K, self._rho

    def continuity_system(self, V_p, V_n, mu=None):
        """(K_n, K_p, f_n, f_p) of the continuity step (device): matrices from the
        current field, mobilities and generation-recombination, then the contact
        values (:847-1035); contact rows replaced."""
        mu_n, mu_p = mu if mu is not None else self._mobilities()
        G = self._generation_recombination()
        _lib.check(_lib.lib.qds_dd_continuity(ctypes.byref(self._m), int(self.ldlt_lower), KT, ptr(self._E),
                                              ptr(mu_n), ptr(mu_p), ptr(G), ptr(self._n), ptr(self._p),
                                              ptr(self._phi_n), ptr(self._phi_p), ptr(self._Kn), ptr(self._Kp),
                                              ptr(self._fn), ptr(self._fp), stream_ptr()), "qds_dd_continuity")
        # contact values after the assembly, as the reference orders it; the replaced
        # rows take the new phi_n / phi_p
        self._contacts(V_p, V_n, loads=True)
        return self._Kn, self._Kp, self._fn, self._fp

    def solve_poisson_equation(self, V_p, V_n):
        """Poisson step (:296-519)."""
        if self.ldlt_lower:
            K, f = self.poisson_system(V_p, V_n)
            x = self._phi.clone()
            self._krylov(K, f, x, "poisson")
            phi = x
        else:
            prob = self._poisson_problem()
            # K phi = -(1/2) M rho: the reference's three-point rule with weights 1/6 on
            # the physical area is half the consistent mass matrix
            _lib.check(_lib.lib.qds_dd_charge(self.n, 0.5 * Q, ptr(self._n), ptr(self._p), ptr(self._doping),
                                              ptr(self._rho), stream_ptr()), "qds_dd_charge")
            phi = prob.solve(self._rho, float(V_p), float(V_n), x0=self._phi, rtol=self.rtol)
            self.last_info["poisson"] = {"iterations": prob.last_iterations, "relres": prob.last_relres}
            if not prob.last_converged:
                self.last_info["unconverged"] = self.last_info.get("unconverged", 0) + 1
        if not bool(torch.isfinite(phi).all()):
            warnings.warn("non-finite Poisson solution; using the linear profile", RuntimeWarning, stacklevel=2)
            phi = V_p + (V_n - V_p) * self._x / self.Lx
        self._phi.copy_(phi)
        self.compute_electric_field()

    def solve_drift_diffusion_equations(self, V_p, V_n):
        """Continuity step (:847-1063)."""
        mu_n, mu_p = self._mobilities()
        self.continuity_system(V_p, V_n, (mu_n, mu_p))
        prob = None if self.ldlt_lower else self._poisson_problem()
        if prob is not None and self._cont_mg is not None:
            # the Dirichlet problems on the free nodes (lattice order), the contact
            # values lifted into the load: K_ff x = f_f - K_fc x_c
            for K, f, x, mg, what in ((self._Kn, self._fn, self._phi_n, self._cont_mg[0], "electrons"),
                                      (self._Kp, self._fp, self._phi_p, self._cont_mg[1], "holes")):
                t = torch.empty_like(x)
                kernels.spmm(self._indptr, self._indices, self.n, self.nnz, K, x * self._mask, t)
                b = (f - t)[prob.free_d].contiguous()
                xf = x[prob.free_d].contiguous()
                K_ff = kernels.csr_take(prob.r_src, K)
                it, rr, rs = kernels.bicgstab_mg(prob.r_indptr, prob.r_indices, K_ff, prob.free.size,
                                                 int(prob.r_indices.numel()), b, xf, mg._structs, mg.n_levels,
                                                 mg.nu, mg.coarse_sweeps, mg.omega, rtol=self.rtol,
                                                 max_iter=self.max_krylov)
                x[prob.free_d] = xf
                self._record(what, it, rr, rs)
        else:
            self._krylov(self._Kn, self._fn, self._phi_n, "electrons")
            self._krylov(self._Kp, self._fp, self._phi_p, "holes")
        self._carriers()
        self.compute_current_densities((mu_n, mu_p))

    def solve(self, V_p, V_n, tolerance=1e-6, max_iter=100, min_iter=0):
        """Gummel loop of :1296-1427 (see the module docstring)."""
        V_p, V_n = float(V_p), float(V_n)
        self.last_info = {}
        self.initialize_carrier_concentrations()
        self._contacts(V_p, V_n)
        self._phi.copy_(V_p + (V_n - V_p) * self._x / self.Lx)
        self.compute_electric_field()
        self._carriers()
        self.compute_electric_field()
        self.compute_current_densities()
        wsb = _lib.lib.qds_dd_workspace()
        error, it = 1.0, 0
        while (error > tolerance or it < min_iter) and it < max_iter:
            self._phi_old.copy_(self._phi)
            if it > 0:
                self.solve_poisson_equation(V_p, V_n)
                if it > 1:
                    _lib.check(_lib.lib.qds_dd_potential_update(self.n, 0, self.damping_factor, ptr(self._phi_old),
                                                                ptr(self._phi), None, None, 0, stream_ptr()),
                               "qds_dd_potential_update")
            self.compute_electric_field()
            self.solve_drift_diffusion_equations(V_p, V_n)
            wp, wn = workspace("dd").ptr(wsb)
            _lib.check(_lib.lib.qds_dd_potential_update(self.n, 1, 0.0, ptr(self._phi_old), ptr(self._phi),
                                                        ptr(self._sums), wp, wn, stream_ptr()),
                       "qds_dd_potential_update")
            pn2, dn2, bad = self._sums[:3].tolist()
            if bad > 0:
                warnings.warn("non-finite potential after the drift-diffusion step; reverted",
                              RuntimeWarning, stacklevel=2)
                error = 1.0
            else:
                pn, dn = math.sqrt(pn2), math.sqrt(dn2)
                error = dn / pn if pn > 1e-10 else dn
            if not math.isfinite(error):
                error = 1.0
            it += 1
        self.iterations, self.error = it, error
        self.compute_electric_field()
        self.compute_current_densities()
        torch.cuda.current_stream().synchronize()

    # -- getters (host copies) -----------------------------------------------------------
    def get_potential(self):
        return self._phi.cpu().numpy()

    def get_electron_concentration(self):
        return self._n.cpu().numpy()

    def get_hole_concentration(self):
        return self._p.cpu().numpy()

    def get_quasi_fermi_levels(self):
        return self._phi_n.cpu().numpy(), self._phi_p.cpu().numpy()

    def get_electric_field_all(self):
        """(num_nodes, 2) array; row i is the reference's E_field[i]."""
        return self._E.cpu().numpy()

    def get_electron_current_density(self):
        return self._Jn.cpu().numpy()

    def get_hole_current_density(self):
        return self._Jp.cpu().numpy()

    def get_electric_field(self, x=None, y=None):
        """Field at a point (:592-651): -grad phi of the element holding (x, y), zero
        outside the mesh; with no point, every node's averaged field."""
        if x is None:
            return self.get_electric_field_all()
        if not (0.0 <= x <= self.Lx and 0.0 <= y <= float(self.mesh.Ly)):
            return np.zeros(2)
        if getattr(self, "_interp", None) is None:
            from .interpolator import FEInterpolator
            self._interp = FEInterpolator(self.mesh)
        _, gx, gy = self._interp.interpolate_points([x], [y], self._phi, gradient=True)
        return -np.array([float(gx[0]), float(gy[0])])

