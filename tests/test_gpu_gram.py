"""GPU tests of the block Gram product G = V^H W (qs_block_gram).

The matrix-core kernel (k_gram_mfma, v_mfma_f64_16x16x4f64 with LDS-staged
operands, FP32 widened to FP64) against the VALU Gram kernel of the Arnoldi
step, one column at a time (the Gram-Schmidt dots of operators.jl:1697-1719),
within 1e-13 relative, and both against numpy's V^H W; then the FP32-stream
eigsh, whose reported eigenvalues come from the FP64 Rayleigh-Ritz products
on this kernel, against the FP64 eigsh on a spectrum with degenerate pairs.
"""
import ctypes

import numpy as np
import pytest

import _dense
import qsgpu
from qsgpu import _lib

pytestmark = pytest.mark.gpu

DT = {np.float64: _lib.QS_F64, np.complex128: _lib.QS_C128, np.float32: _lib.QS_F32, np.complex64: _lib.QS_C64}


def _gram(V, W, mfma):
    n, m = V.shape
    p = W.shape[1]
    G = np.zeros(2 * m * p)
    Vf, Wf = np.asfortranarray(V), np.asfortranarray(W)
    qsgpu.errors.check(_lib.lib().qs_block_gram(DT[V.dtype.type], Vf.ctypes.data, n, m, Wf.ctypes.data, p,
                                                int(mfma), G.ctypes.data))
    return (G[0::2] + 1j * G[1::2]).reshape((m, p), order="F")


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.complex64])
@pytest.mark.parametrize("shape", [(5, 2, 1), (1000, 4, 4), (100003, 20, 3), (70001, 33, 17), (4096, 16, 16),
                                   (30011, 7, 40), (3001, 48, 50)])
def test_block_gram_mfma_against_valu_and_numpy(gpu, dtype, shape):
    n, m, p = shape
    print("block gram", np.dtype(dtype).name, shape, flush=True)
    rng = np.random.default_rng(n + m + p)

    def mat(c):
        a = rng.standard_normal((n, c))
        if np.issubdtype(dtype, np.complexfloating):
            a = a + 1j * rng.standard_normal((n, c))
        return a.astype(dtype)

    V, W = mat(m), mat(p)
    g_mfma = _gram(V, W, True)
    g_valu = _gram(V, W, False)
    exact = V.astype(np.complex128).conj().T @ W.astype(np.complex128)
    scale = np.sqrt(n) * max(1.0, np.abs(exact).max() / np.sqrt(n))
    assert np.abs(g_mfma - g_valu).max() <= 1e-13 * scale
    assert np.abs(g_mfma - exact).max() <= 1e-12 * scale
    if not np.issubdtype(dtype, np.complexfloating):
        assert np.all(g_mfma.imag == 0)


def test_block_gram_rejects_bad_shapes(gpu):
    V = np.zeros((10, 2))
    G = np.zeros(8)
    st = _lib.lib().qs_block_gram(_lib.QS_F64, V.ctypes.data, 10, 0, V.ctypes.data, 2, 1, G.ctypes.data)
    assert st == _lib.QS_EARG


def test_fp32_eigsh_rayleigh_ritz_on_degenerate_pairs(gpu):
    # L=16 nup=8 XXZ without symmetry blocks: the low spectrum holds degenerate
    # pairs (momenta +-k); FP32-stored basis, eigenvalues from the FP64
    # Rayleigh-Ritz step, against the FP64 run within 1e-8
    print("fp32 eigsh degenerate pairs", flush=True)
    L = 16
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, _dense.xxz_terms(L, J=1.0, delta=0.5))
    v0 = np.random.default_rng(9).standard_normal(b.Ns)
    e64 = qsgpu.eigsh(op, k=5, which="SA", tol=1e-10, v0=v0, return_eigenvectors=False)
    e32, x32 = qsgpu.eigsh(op, k=5, which="SA", tol=1e-10, v0=v0, dtype=np.float32)
    assert np.min(np.diff(e64)) < 1e-9, e64        # the spectrum does hold a repeated value
    assert np.abs(e32 - e64).max() <= 1e-8, (e32, e64)
    assert x32.dtype == np.float32
    for i in range(5):
        x = x32[:, i].astype(np.float64)
        r = np.linalg.norm(op.matvec(x) - e32[i] * x) / np.linalg.norm(x)
        assert r <= 1e-5, (i, r)
