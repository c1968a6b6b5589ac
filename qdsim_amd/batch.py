"""Batched drivers for sweeps of independent problems: bias / geometry sweeps of
the self-consistent Poisson-Schroedinger loop and parameter sets of the
differentiable eigenvalues.

The instances of a sweep never communicate, so a batch runs them on a small pool
of host threads, each with its own HIP stream (default priority) and its own
scratch workspaces, or in worker processes (``run_processes``), each with its
own interpreter and HIP context, so that the instances' host loops do not share
one interpreter lock.  Every instance synchronises with the host once per block
iteration (Rayleigh-Ritz on the host, convergence test); while one waits, the
others' kernels keep the GPU busy, and small meshes, whose kernels are too short
to fill the chip, overlap with each other.  An instance gives the same result
as when it runs alone: same kernels, same seeded starts, single-threaded host
LAPACK.

Reference: the self-consistent loop of
qdsim_cython/solvers/self_consistent_solver.pyx:130-301 runs one bias point per
call; the sweep and gradient drivers have no reference counterpart.
"""
import threading

import numpy as np
import torch
from threadpoolctl import threadpool_limits

from .device import device, sync
from .units import EV_TO_J

M_E = 9.1093837015e-31


def run_batch(fn, items, workers=4, setup=None, progress=None):
    """[fn(state, item) for item in items] on ``workers`` host threads, each with
    its own stream; ``state = setup()`` is built once per thread (None without
    setup).  Items are taken in order from a shared queue; the results keep the
    order of ``items``.  The first exception stops the batch and is re-raised.
    progress(done, total), if given, is called as each item finishes."""
    items = list(items)
    results = [None] * len(items)
    if not items:
        return results
    dev = device()
    lock = threading.Lock()
    nxt = [0]
    done = [0]
    errors = []

    def worker():
        stream = torch.cuda.Stream(device=dev)
        try:
            with torch.cuda.stream(stream):
                state = setup() if setup is not None else None
                while True:
                    with lock:
                        if errors or nxt[0] >= len(items):
                            break
                        i = nxt[0]
                        nxt[0] += 1
                    results[i] = fn(state, items[i])
                    if progress is not None:
                        with lock:
                            done[0] += 1
                            progress(done[0], len(items))
        except BaseException as e:   # noqa: BLE001 -- re-raised in the caller's thread
            with lock:
                errors.append(e)
        finally:
            stream.synchronize()

    with threadpool_limits(1, "blas"):
        n = max(1, min(int(workers), len(items)))
        if n == 1:
            worker()
        else:
            threads = [threading.Thread(target=worker, name=f"qdsim-batch-{t}") for t in range(n)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
    sync()
    if errors:
        raise errors[0]
    return results


# ---------------------------------------------------------------- bias sweep
def bias_sweep_instances(n=256, seed=4, cells=(64, 384), orders=(1, 2), Lx=200e-9, Ly=100e-9):
    """Instances of a p-n junction bias / geometry sweep: Lx x Ly domain, c x c
    cells with c uniform in ``cells`` (a multiple of 8), P1 or P2, abrupt junction
    at Lx/2 + U(-10, 10) nm (acceptors left, donors right, 1e24 m^-3), a Gaussian
    dot V(r) = -V0 exp(-r^2 / 2 (10 nm)^2) at the centre with V0 ~ U(0.1, 0.4) eV,
    bias V_n - V_p ~ U(-0.5, 0.5) V, eps_r = 13.9, m* = 0.067 m0."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = int(rng.integers(cells[0] // 8, cells[1] // 8 + 1)) * 8
        out.append({
            "index": i,
            "cells": c,
            "order": int(rng.choice(orders)),
            "junction": Lx / 2 + float(rng.uniform(-10e-9, 10e-9)),
            "V0": float(rng.uniform(0.1, 0.4)),
            "bias": float(rng.uniform(-0.5, 0.5)),
            "Lx": Lx, "Ly": Ly, "N_A": 1e24, "N_D": 1e24, "eps_r": 13.9, "m_star": 0.067,
            "sigma_dot": 10e-9,
        })
    return out


def _sc_instance(spec, num_states, iterations, mixing, keep_fields, poisson_sign="physical", retry=True):
    from .mesh import SimpleMesh
    from .self_consistent import SelfConsistentSolver
    Lx, Ly, c = spec["Lx"], spec["Ly"], spec["cells"]
    mesh = SimpleMesh(c + 1, c + 1, Lx, Ly, spec["order"])
    x0, y0, s2 = Lx / 2, Ly / 2, 2.0 * spec["sigma_dot"] ** 2
    V0 = spec["V0"] * EV_TO_J
    m = spec["m_star"] * M_E
    sc = SelfConsistentSolver(mesh, lambda x, y: spec["eps_r"] + 0 * x, lambda x, y: m + 0 * x,
                              lambda x, y: -V0 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / s2),
                              mixing=mixing, poisson_sign=poisson_sign)
    sc.schrodinger.retry = retry
    sc.set_solver_parameters(max_iterations=iterations, tolerance=0.0)
    x, _ = mesh.get_nodes()
    left = x < spec["junction"]
    res = sc.solve_self_consistent(0.0, spec["bias"], num_quantum_states=num_states,
                                   N_A=np.where(left, spec["N_A"], 0.0), N_D=np.where(left, 0.0, spec["N_D"]))
    out = {"index": spec.get("index"), "cells": c, "order": spec["order"], "dofs": mesh.num_nodes,
           "eigenvalues_eV": res["eigenvalues_eV"], "iterations": res["iterations"],
           "convergence_history": res["convergence_history"], "solve_time": res["solve_time"],
           "breakdown_solves": sc.breakdown_solves, "retried_solves": sc.retried_solves,
           "unconverged_solves": sc.unconverged_solves}
    if keep_fields:
        out.update({k: res[k] for k in ("potential", "electron_density", "hole_density")})
    return out


def _process_init():
    # a worker process of a sweep: its own interpreter and, at its first kernel,
    # its own HIP context and queues on the current GPU; host LAPACK, OpenMP and
    # torch's CPU ops on one thread each (the workers already run side by side)
    threadpool_limits(1)
    torch.set_num_threads(1)


def _sc_instance_task(args):
    spec, num_states, iterations, mixing, keep_fields, poisson_sign, retry = args
    out = _sc_instance(spec, num_states, iterations, mixing, keep_fields, poisson_sign, retry)
    sync()
    return out


def run_processes(fn, items, processes, progress=None):
    """[fn(item) for item in items] in ``processes`` worker processes (spawned: fresh
    interpreters, each with its own HIP context on the same GPU), one item at a
    time, results in the order of ``items``.  Unlike threads, the workers' host
    work (Python per block iteration) does not share one interpreter lock, and the
    GPU runs their kernels concurrently from separate queues.  fn must be a
    module-level function; its exceptions are re-raised here.  The workers are
    started as new interpreters (fork + exec, as subprocess does), never as forks
    of a process that may hold a HIP context.  progress(done, total), if given,
    is called in this process as each item finishes."""
    import concurrent.futures
    import multiprocessing
    items = list(items)
    if not items:
        return []
    ctx = multiprocessing.get_context("spawn")
    n = max(1, min(int(processes), len(items)))
    results = [None] * len(items)
    with concurrent.futures.ProcessPoolExecutor(max_workers=n, mp_context=ctx, initializer=_process_init) as ex:
        futures = {ex.submit(fn, item): i for i, item in enumerate(items)}
        try:
            for done, f in enumerate(concurrent.futures.as_completed(futures), 1):
                results[futures[f]] = f.result()
                if progress is not None:
                    progress(done, len(items))
        except BaseException:
            for f in futures:
                f.cancel()
            raise
    return results


def bias_sweep(instances, num_quantum_states=8, iterations=30, mixing=0.3, workers=4, keep_fields=False,
               poisson_sign="physical", retry=True, processes=0, progress=None):
    """Self-consistent loop of every instance (``bias_sweep_instances``): exactly
    ``iterations`` outer iterations with linear mixing, CG Poisson and the lowest
    ``num_quantum_states`` states.  Returns one summary dict per instance (fields
    on the host only with keep_fields), with the number of eigensolves that broke
    down or did not converge at first (``breakdown_solves``), were redone with the
    safe shift (``retried_solves``; with retry=False a breakdown raises instead and
    a solve that does not converge is kept) and ended unconverged.  poisson_sign: "physical" or "reference" (the
    reference's K phi = -M rho).  processes > 0 runs the instances in that many
    worker processes (``run_processes``) instead of ``workers`` threads;
    progress(done, total) is called as each instance ends.  Instances start largest mesh first, so that
    no large one is left to run alone at the end; the results keep the order of
    ``instances``."""
    instances = list(instances)
    order = sorted(range(len(instances)),
                   key=lambda i: -(instances[i]["cells"] * instances[i]["order"]) ** 2)
    if processes and processes > 0:
        out = run_processes(_sc_instance_task, [(instances[i], num_quantum_states, iterations, mixing,
                                                 keep_fields, poisson_sign, retry) for i in order],
                            processes, progress)
    else:
        out = run_batch(lambda _, spec: _sc_instance(spec, num_quantum_states, iterations, mixing, keep_fields,
                                                     poisson_sign, retry),
                        [instances[i] for i in order], workers=workers, progress=progress)
    results = [None] * len(instances)
    for j, i in enumerate(order):
        results[i] = out[j]
    return results


# ---------------------------------------------------------------- gate gradients
def gate_parameter_sets(n=64, seed=5, n_gates=8, L=100.0, k=8):
    """Parameter sets of the gate-design objective: per set, n_gates Gaussian bumps
    with amplitude a ~ U(-0.3, 0.1) eV, centre uniform in the L x L nm domain and
    width sigma ~ U(4, 12) nm; theta = (a, cx, cy) per gate (3 n_gates
    parameters), sigma fixed per set.  Targets: k energies from the seed, sorted,
    U(-0.25, 0.0) eV.  Returns (theta (n, 3 n_gates), sigma (n, n_gates), targets (k,))."""
    rng = np.random.default_rng(seed)
    theta = np.empty((n, 3 * n_gates))
    theta[:, 0::3] = rng.uniform(-0.3, 0.1, (n, n_gates))
    theta[:, 1::3] = rng.uniform(0.0, L, (n, n_gates))
    theta[:, 2::3] = rng.uniform(0.0, L, (n, n_gates))
    sigma = rng.uniform(4.0, 12.0, (n, n_gates))
    targets = np.sort(rng.uniform(-0.25, 0.0, k))
    return theta, sigma, targets


def gate_potential(theta, sigma, xq, yq):
    """V (eV) at the quadrature points (nm): sum_g a_g exp(-|r - c_g|^2 / 2 sigma_g^2),
    a differentiable torch expression of theta."""
    v = torch.zeros_like(xq)
    for g in range(sigma.shape[0]):
        a, cx, cy = theta[3 * g], theta[3 * g + 1], theta[3 * g + 2]
        v = v + a * torch.exp(-((xq - cx) ** 2 + (yq - cy) ** 2) / (2.0 * float(sigma[g]) ** 2))
    return v


def gate_gradients(theta, sigma, targets, cells=256, order=2, L=100e-9, m_star=0.067, mode="mixed", tol=1e-9,
                   workers=1):
    """Loss sum_k (E_k - E*_k)^2 over the lowest len(targets) states and its gradient
    with respect to every parameter set's theta, through the Hellmann-Feynman
    adjoint (``adjoint.eigenvalues``).  Each worker builds the problem once and
    reuses it for its sets.  One worker by default: on one mesh of 263k DOF the
    solves are short and host-bound, and concurrent workers only contend for the
    interpreter (16 sets: 0.67 s with one worker, 0.96 s with four).  Returns
    (loss (n,), grad (n, p), E (n, k)) as numpy."""
    from .adjoint import eigenvalues
    from .mesh import SimpleMesh
    from .schrodinger import SchrodingerProblem
    k = len(targets)
    m = m_star * M_E

    def setup():
        prob = SchrodingerProblem(SimpleMesh(cells + 1, cells + 1, L, L, order), lambda x, y: m + 0 * x, None)
        xq, yq = prob.space.quad_points()
        return prob, xq, yq

    def one(state, i):
        prob, xq, yq = state
        th = torch.tensor(theta[i], dtype=torch.float64, device=xq.device, requires_grad=True)
        E = eigenvalues(prob, gate_potential(th, sigma[i], xq, yq), k, mode=mode, tol=tol)
        loss = ((E - torch.as_tensor(targets, dtype=torch.float64, device=E.device)) ** 2).sum()
        (g,) = torch.autograd.grad(loss, th)
        return float(loss.item()), g.cpu().numpy(), E.detach().cpu().numpy()

    out = run_batch(one, range(theta.shape[0]), workers=workers, setup=setup)
    return (np.array([o[0] for o in out]), np.stack([o[1] for o in out]), np.stack([o[2] for o in out]))


def gate_potential_batched(theta, sigma, xq, yq):
    """V (eV) of every parameter set at the quadrature points: theta (B, 3 G) torch
    (differentiable), sigma (B, G) numpy; returns (B, E, Q)."""
    sig = torch.as_tensor(np.asarray(sigma, dtype=np.float64), device=theta.device)
    a = theta[:, 0::3][:, :, None, None]
    cx = theta[:, 1::3][:, :, None, None]
    cy = theta[:, 2::3][:, :, None, None]
    s2 = (2.0 * sig ** 2)[:, :, None, None]
    return (a * torch.exp(-((xq[None, None] - cx) ** 2 + (yq[None, None] - cy) ** 2) / s2)).sum(dim=1)


class GateBatchSolver:
    """All parameter sets of the gate-design objective on one mesh, solved together
    (batched_eigen.BatchedSchrodinger): one launch per solver stage for the whole
    batch.  The gradient of sum_k (E_k - E*_k)^2 with respect to every set's theta is
    Hellmann-Feynman (dE_k/dV at the quadrature points, qds_state_density_q) chained
    through the batched gate expression by one autograd call."""

    def __init__(self, cells=256, order=2, L=100e-9, m_star=0.067):
        from .batched_eigen import BatchedSchrodinger
        from .mesh import SimpleMesh
        m = m_star * M_E
        self.bs = BatchedSchrodinger(SimpleMesh(cells + 1, cells + 1, L, L, order), lambda x, y: m + 0 * x)
        self.xq, self.yq = self.bs.quad_points_nm()

    def __call__(self, theta, sigma, targets, tol=1e-9, verbose=False):
        from .adjoint import state_density_q
        dev = self.xq.device
        th = torch.as_tensor(np.asarray(theta, dtype=np.float64), device=dev).requires_grad_(True)
        tgt = torch.as_tensor(np.asarray(targets, dtype=np.float64), device=dev)
        k = len(targets)
        v_q = gate_potential_batched(th, sigma, self.xq, self.yq)
        v_det = v_q.detach()

        def vfun(xq, yq):
            if xq.shape == self.xq.shape and xq.data_ptr() == self.xq.data_ptr():
                return v_det
            with torch.no_grad():
                return gate_potential_batched(th.detach(), sigma, xq, yq)
        lam, psi, info = self.bs.solve(vfun, k, tol=tol, verbose=verbose)
        E = torch.as_tensor(lam, device=dev)
        gE = 2.0 * (E - tgt[None])
        gV = torch.empty_like(v_det)
        for i in range(E.shape[0]):
            rho = state_density_q(self.bs.space, psi[i])          # (E, Q, k)
            gV[i] = rho @ gE[i]
        (g,) = torch.autograd.grad(v_q, th, grad_outputs=gV)
        loss = ((E - tgt[None]) ** 2).sum(dim=1)
        self.last_info = info
        return loss.cpu().numpy(), g.cpu().numpy(), lam


def gate_gradients_batched(theta, sigma, targets, cells=256, order=2, L=100e-9, m_star=0.067, tol=1e-9,
                           solver=None):
    """gate_gradients for all parameter sets in one batched solve.  Returns (loss (n,),
    grad (n, p), E (n, k)) as numpy; ``solver``: a GateBatchSolver to reuse."""
    solver = solver or GateBatchSolver(cells, order, L, m_star)
    return solver(theta, sigma, targets, tol=tol)
