"""GPU tests of the single-workgroup Krylov path for small assembled sectors.

An assembled operator of at most 2048 rows runs each thick-restart Lanczos
cycle of eigsh, and each expm factorization, in one launch of one workgroup
(k_krylov_small); the same operator left matrix-free runs the per-operation
launches.  Both follow the same step sequence (operators.jl:1655-1731 for the
factorization, ARPACK's DGKS step for the cycle) and differ only in the order
of their FP64 sums, so eigenvalues agree to 1e-10 and evolved states to 1e-12,
for the three step forms: DGKS cycle, three-term Lanczos (Hermitian, FP64)
and two-pass Arnoldi (FP32 storage, and a non-Hermitian operator), against
the oracle's dense sector matrix as well.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle
import qsgpu
from qsgpu import serve

pytestmark = pytest.mark.gpu


def _pair(L, terms, **blocks):
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False, **blocks)
    free = qsgpu.QuantumLinearOperator(b, terms)
    asm = qsgpu.QuantumLinearOperator(b, terms).assemble()
    return b, free, asm


@pytest.mark.parametrize("L,blocks", [(12, dict(kblock=0, pblock=1)), (14, dict(kblock=3)), (16, dict(kblock=1))])
def test_small_eigsh_matches_per_operation_path(gpu, L, blocks):
    print("small eigsh", L, blocks, flush=True)
    terms = serve.chain_terms(L, 0.3, 0.9)
    b, free, asm = _pair(L, terms, **blocks)
    assert b.Ns <= 2048
    v0 = np.random.default_rng(L).standard_normal(b.Ns) + 0j
    for k in (1, 4):
        e_free = qsgpu.eigsh(free, k=k, which="SA", tol=1e-12, v0=v0, return_eigenvectors=False)
        e_asm = qsgpu.eigsh(asm, k=k, which="SA", tol=1e-12, v0=v0, return_eigenvectors=False)
        assert np.abs(np.sort(e_free) - np.sort(e_asm)).max() <= 1e-10, (e_free, e_asm)
    sec = oracle.Sector(L, nup=L // 2, pauli=False, **blocks)
    w = np.linalg.eigvalsh(oracle.Operator(sec, terms).dense())
    assert np.abs(np.sort(e_asm) - w[:4]).max() <= 1e-10


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_small_expm_matches_per_operation_path(gpu, dtype):
    print("small expm", np.dtype(dtype).name, flush=True)
    L = 14
    terms = serve.chain_terms(L, 0.2, 1.3)
    b, free, asm = _pair(L, terms, kblock=2)
    v = np.random.default_rng(3).standard_normal(b.Ns) + 1j * np.random.default_rng(4).standard_normal(b.Ns)
    v /= np.linalg.norm(v)
    sc = np.array([-0.7j, -3.0j, -0.4])
    tol = 1e-12 if dtype == np.complex128 else 1e-6
    y_free = qsgpu.expm_multiply(free, v, sc, tol=tol, dtype=dtype)
    y_asm = qsgpu.expm_multiply(asm, v, sc, tol=tol, dtype=dtype)
    H = oracle.Operator(oracle.Sector(L, nup=L // 2, pauli=False, kblock=2), terms).dense()
    want = np.stack([scipy.linalg.expm(s * H) @ v for s in sc], axis=-1)
    bar = 1e-11 if dtype == np.complex128 else 2e-5
    assert np.abs(y_asm - y_free).max() <= bar * np.abs(want).max()
    assert np.abs(y_asm - want).max() <= bar * np.abs(want).max()


def test_small_non_hermitian_expm(gpu):
    # a non-Hermitian coupling list takes the two-pass Arnoldi form
    print("small expm non-Hermitian", flush=True)
    L = 10
    terms = [("+-", [(0.8, i, i % L + 1) for i in range(1, L + 1)]),
             ("-+", [(0.3, i, i % L + 1) for i in range(1, L + 1)]),
             ("zz", [(0.5, i, i % L + 1) for i in range(1, L + 1)])]
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False)
    free = qsgpu.QuantumLinearOperator(b, terms, check_herm=False)
    asm = qsgpu.QuantumLinearOperator(b, terms, check_herm=False).assemble()
    v = np.random.default_rng(8).standard_normal(b.Ns) + 0j
    sc = np.array([-0.5j, -2.0j])
    y_free = qsgpu.expm_multiply(free, v, sc)
    y_asm = qsgpu.expm_multiply(asm, v, sc)
    H = oracle.Operator(oracle.Sector(L, nup=L // 2, pauli=False), terms, check_symm=False).dense()
    want = np.stack([scipy.linalg.expm(s * H) @ v for s in sc], axis=-1)
    assert np.abs(y_asm - y_free).max() <= 1e-11 * np.abs(want).max()
    assert np.abs(y_asm - want).max() <= 1e-11 * np.abs(want).max()
