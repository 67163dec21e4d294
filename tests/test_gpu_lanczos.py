"""GPU Lanczos tools (tools.jl:2742-2952) against the reference's own known
answers (runtests.jl:131-154, committed in tests/golden) and against
orthogonality / projection identities on a spin sector."""
import json
import os

import numpy as np
import pytest

import qsgpu

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


def _two_site_operator(A):
    # the 4 x 4 matrix as a spin-1/2 operator on L=2 (no y: every piece is real)
    b = qsgpu.SpinBasis1D(2, pauli=True)
    pieces = [(a + c, (1, 2)) for a in "Iz+-" for c in "Iz+-"]
    mats = []
    for op_str, sites in pieces:
        mats.append(qsgpu.QuantumLinearOperator(b, [(op_str, [(1.0,) + sites])], check_herm=False,
                                                check_symm=False, check_pcon=False).toarray().real)
    M = np.stack([m.ravel() for m in mats], axis=1)
    coef, *_ = np.linalg.lstsq(M, np.asarray(A).ravel(), rcond=None)
    terms = [(op_str, [(float(c),) + sites]) for (op_str, sites), c in zip(pieces, coef) if abs(c) > 1e-15]
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False, check_pcon=False)
    assert np.abs(op.toarray() - np.asarray(A)).max() <= 1e-15
    return op


def test_lanczos_full_known_answers(gpu):
    g = GOLD["lanczos_full_4x4"]
    op = _two_site_operator(g["A"])
    E, V, Q_T = qsgpu.lanczos_full(op, np.array(g["v0"]), g["m"])
    assert np.allclose(E, g["E"], atol=5e-15, rtol=0)
    assert np.abs(Q_T @ Q_T.conj().T - np.eye(3)).max() <= 2e-15
    assert np.abs(qsgpu.lin_comb_Q_T(g["lin_comb_coeff"], Q_T) - g["lin_comb"]).max() <= 2e-15
    assert np.abs(qsgpu.expm_lanczos(E, V, Q_T, a=g["expm_lanczos_a"]) - g["expm_lanczos"]).max() <= 5e-15
    E2, V2, it = qsgpu.lanczos_iter(op, np.array(g["v0"]), g["m"])
    assert np.allclose(E2, E, atol=1e-14, rtol=0)
    assert np.allclose(np.abs(V2), np.abs(V), atol=1e-14, rtol=0)
    assert np.abs(np.stack(list(it)) - Q_T).max() <= 1e-14
    assert np.abs(qsgpu.lin_comb_Q_T(g["lin_comb_coeff"], iter(qsgpu.lanczos_iter(op, np.array(g["v0"]), 3)[2]))
                  - g["lin_comb"]).max() <= 2e-15


def test_lanczos_on_a_sector(gpu):
    # Q_T H Q_T^dagger is the SymTridiagonal of the recurrence; full_ortho keeps Q_T orthonormal
    L = 12
    b = qsgpu.SpinBasis1D(L, nup=6, pauli=False, kblock=0, pblock=1)
    bonds = [(i, i % L + 1) for i in range(1, L + 1)]
    op = qsgpu.QuantumLinearOperator(b, [("+-", [(0.5, i, j) for i, j in bonds]),
                                         ("-+", [(0.5, i, j) for i, j in bonds]),
                                         ("zz", [(0.8, i, j) for i, j in bonds])])
    H = op.toarray()
    v0 = np.random.default_rng(4).standard_normal(b.Ns)
    E, V, Q = qsgpu.lanczos_full(op, v0, 20, full_ortho=True)
    assert np.abs(Q @ Q.T - np.eye(Q.shape[0])).max() <= 1e-12
    T = Q @ H @ Q.T
    assert np.allclose(np.linalg.eigvalsh(T), E, atol=1e-10, rtol=0)
    assert abs(E[0] - np.linalg.eigvalsh(H)[0]) <= 1e-8
    with pytest.raises(qsgpu.ArgumentError):
        qsgpu.lanczos_full(op, v0, b.Ns)
