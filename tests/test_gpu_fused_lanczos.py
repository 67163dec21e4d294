"""GPU tests of the fused Lanczos product of an assembled operator.

Above the single-workgroup size (2048 rows) a Hermitian expm factorization of
an assembled operator runs each three-term Lanczos step as one product that
also subtracts beta v_{j-1} and leaves the per-block partials of <v_j, w>
(operators.jl:1655-1731, the same operation order as the separate product,
axpby and dot); a matrix-free operator runs the separate launches.  Both are
checked against the numpy restatement of the Krylov exponential action
(oracle/krylov.py) on the oracle's sector operator, and against each other.
"""
import numpy as np
import pytest

import oracle
from oracle import krylov
import qsgpu
from qsgpu import serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,blocks", [(20, dict(kblock=2)), (20, dict(kblock=0, pblock=1, zblock=-1))])
def test_fused_lanczos_expm_matches_restatement(gpu, L, blocks):
    print("fused lanczos expm", L, blocks, flush=True)
    terms = serve.chain_terms(L, 0.35, 0.8)
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False, **blocks)
    assert b.Ns > 2048
    free = qsgpu.QuantumLinearOperator(b, terms)
    asm = qsgpu.QuantumLinearOperator(b, terms).assemble()
    rng = np.random.default_rng(L + b.Ns)
    v = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    v /= np.linalg.norm(v)
    sc = np.array([-0.3j, -2.5j, -7.0j])
    y_asm = qsgpu.expm_multiply(asm, v, sc, tol=1e-12)
    y_free = qsgpu.expm_multiply(free, v, sc, tol=1e-12)
    H = oracle.Operator(oracle.Sector(L, nup=L // 2, pauli=False, **blocks), terms)
    want = krylov.expmv_times(lambda x: H.matvec(x), v, sc, hermitian=True)
    assert np.abs(y_asm - want).max() <= 1e-10, np.abs(y_asm - want).max()
    assert np.abs(y_asm - y_free).max() <= 1e-10
    # unitary evolution keeps the norm
    assert np.abs(np.linalg.norm(y_asm, axis=0) - 1).max() <= 1e-10


def test_fused_lanczos_evolve_in_place_of_separate_steps(gpu):
    # the same assembled operator with its product forced matrix-free gives
    # the separate-launch form of the same recurrence
    print("fused lanczos vs separate on one operator", flush=True)
    L = 22
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False, kblock=3)
    assert b.Ns > 2048
    op = qsgpu.QuantumLinearOperator(b, serve.chain_terms(L, 0.2, 1.5)).assemble()
    v = np.random.default_rng(9).standard_normal(b.Ns) + 0j
    v /= np.linalg.norm(v)
    y_fused = qsgpu.expm_multiply(op, v, np.array([-4.0j]), tol=1e-12)
    op.set_variant(1)   # the matrix-free pull kernel, one lane per row
    y_sep = qsgpu.expm_multiply(op, v, np.array([-4.0j]), tol=1e-12)
    assert np.abs(y_fused - y_sep).max() <= 1e-11
