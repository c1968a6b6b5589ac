"""Batched sweeps (bias sweep of the self-consistent loop, gate-parameter gradients):
instances run concurrently on their own streams give the results of running them
one after another, and the batched gradients match central differences."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_bias_sweep_concurrent_equals_sequential():
    from qdsim_amd.batch import bias_sweep, bias_sweep_instances
    inst = bias_sweep_instances(n=4, seed=4, cells=(16, 32), orders=(1, 2))
    print("bias sweep instances", [(s["cells"], s["order"], round(s["bias"], 3)) for s in inst], flush=True)
    seq = bias_sweep(inst, num_quantum_states=4, iterations=3, workers=1, keep_fields=True)
    par = bias_sweep(inst, num_quantum_states=4, iterations=3, workers=3, keep_fields=True)
    for a, b in zip(seq, par):
        assert a["iterations"] == b["iterations"] == 3
        assert len(a["eigenvalues_eV"]) == 4
        assert np.allclose(a["eigenvalues_eV"], b["eigenvalues_eV"], rtol=1e-12, atol=0)
        assert np.allclose(a["potential"], b["potential"], rtol=0, atol=1e-12)
        assert np.allclose(a["convergence_history"], b["convergence_history"], rtol=1e-9, atol=1e-15)
    # the junction and the bias enter: potentials differ between instances
    assert not np.allclose(seq[0]["potential"][:10], seq[1]["potential"][:10])


def test_bias_sweep_processes_equal_threads():
    # worker processes (own interpreter and HIP context each) give the thread
    # workers' results: same kernels, same seeded starts (each process tunes its
    # own launch shapes, so compare to FP64 rounding, not bits)
    from qdsim_amd.batch import bias_sweep, bias_sweep_instances
    inst = bias_sweep_instances(n=3, seed=6, cells=(16, 32), orders=(1, 2))
    print("bias sweep instances (processes)", [(s["cells"], s["order"]) for s in inst], flush=True)
    thr = bias_sweep(inst, num_quantum_states=4, iterations=2, workers=2, keep_fields=True)
    prc = bias_sweep(inst, num_quantum_states=4, iterations=2, keep_fields=True, processes=2)
    assert [o["index"] for o in prc] == [s["index"] for s in inst]
    for a, b in zip(thr, prc):
        assert a["iterations"] == b["iterations"] == 2
        assert np.allclose(a["eigenvalues_eV"], b["eigenvalues_eV"], rtol=1e-12, atol=0)
        assert np.allclose(a["potential"], b["potential"], rtol=0, atol=1e-12)


def test_gate_gradients_concurrent_equal_sequential_and_fd():
    from qdsim_amd.batch import gate_gradients, gate_parameter_sets
    theta, sigma, targets = gate_parameter_sets(n=3, seed=5, n_gates=2, L=40.0, k=3)
    kw = dict(cells=12, order=2, L=40e-9, mode="fp64", tol=1e-10)
    print("gate sets", theta.round(2).tolist(), flush=True)
    l1, g1, e1 = gate_gradients(theta, sigma, targets, workers=1, **kw)
    l3, g3, e3 = gate_gradients(theta, sigma, targets, workers=3, **kw)
    assert g1.shape == (3, 6) and e1.shape == (3, 3)
    assert np.allclose(g1, g3, rtol=1e-10, atol=1e-14) and np.allclose(l1, l3, rtol=1e-12)
    # central differences (step 1e-6) on set 1
    h = 1e-6
    for j in (0, 1, 5):
        tp = theta.copy(); tp[1, j] += h
        tm = theta.copy(); tm[1, j] -= h
        lp = gate_gradients(tp[1:2], sigma[1:2], targets, workers=1, **kw)[0][0]
        lm = gate_gradients(tm[1:2], sigma[1:2], targets, workers=1, **kw)[0][0]
        fd = (lp - lm) / (2 * h)
        print("param", j, "adjoint", g1[1, j], "fd", fd)
        assert abs(g1[1, j] - fd) <= 1e-4 * abs(fd) + 1e-7
