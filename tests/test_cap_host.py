"""Host-side algebra of the CAP eigensolver (CPU): the interleaved real views of
complex blocks, the bilinear / Hermitian forms read from their real Gram, the
2x2-real form of a complex coefficient matrix, and the complex-symmetric
Cholesky factor used to orthonormalise a deflation basis."""
import numpy as np
import torch

from qdsim_amd.cap import _bilinear, _csym_chol, _hermitian, _real_form, _rv


def _cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_real_gram_gives_bilinear_and_hermitian_products():
    rng = np.random.default_rng(0)
    S, B = _cplx(rng, 50, 3), _cplx(rng, 50, 5)
    Sr = _rv(torch.from_numpy(S)).numpy()
    Br = _rv(torch.from_numpy(B)).numpy()
    G = Sr.T @ Br
    assert np.allclose(_bilinear(G), S.T @ B, rtol=0, atol=1e-12)
    assert np.allclose(_hermitian(G), S.conj().T @ B, rtol=0, atol=1e-12)


def test_real_form_applies_complex_coefficients_to_interleaved_columns():
    rng = np.random.default_rng(1)
    S, C = _cplx(rng, 40, 4), _cplx(rng, 4, 3)
    Sr = _rv(torch.from_numpy(S))
    Y = Sr @ _real_form(torch.from_numpy(C))
    assert np.allclose(Y.numpy(), _rv(torch.from_numpy(S @ C)).numpy(), rtol=0, atol=1e-12)


def test_complex_symmetric_cholesky():
    rng = np.random.default_rng(2)
    E = 0.1 * _cplx(rng, 6, 6)
    G = np.eye(6) * np.exp(0.3j) + E + E.T          # complex symmetric, not Hermitian
    L = _csym_chol(G)
    assert np.allclose(np.triu(L, 1), 0)
    assert np.allclose(L @ L.T, G, rtol=0, atol=1e-12)
    # the orthonormalised basis Q L^-T has bilinear Gram I
    Li = np.linalg.inv(L)
    assert np.allclose(Li @ G @ Li.T, np.eye(6), rtol=0, atol=1e-12)
