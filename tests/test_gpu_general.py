"""GPU tests of SpinBasisGeneral (basis.jl:530-620): sectors of arbitrary site maps.

Each sector is built twice: by libqsgpu (the whole group's orbit minima on the
device, the parent lookup table, the reference's labels and normalizations
from its stage sequence) and by the oracle's restatement of the reference's
stage-by-stage projector for the same maps in the same block order.  Bars:
states, orbit sizes and normalizations bit-exact; the sector Hamiltonian's
product within 1e-12 of the oracle's lift / apply / project; eigenvalues
within 1e-10 of the oracle's dense sector matrix; the reference's argument
errors.
"""
import numpy as np
import pytest

import oracle
import qsgpu

pytestmark = pytest.mark.gpu


def _translate(L, d):
    return [(i + d) % L for i in range(L)]            # zero-based labels


def _reflect(L):
    return [L - 1 - i for i in range(L)]


def _invert(L):
    return [-(i + 1) for i in range(L)]               # every site inverted, to itself


def _pair_swap(L):
    return [i + 1 if i % 2 == 0 else i - 1 for i in range(L)]


def _terms(L):
    b1 = [(i, i % L + 1) for i in range(1, L + 1)]
    b2 = [(i, (i + 1) % L + 1) for i in range(1, L + 1)]
    return [("+-", [(0.5, i, j) for i, j in b1]), ("-+", [(0.5, i, j) for i, j in b1]),
            ("zz", [(0.8, i, j) for i, j in b1]),
            ("+-", [(0.2, i, j) for i, j in b2]), ("-+", [(0.2, i, j) for i, j in b2]),
            ("zz", [(0.3, i, j) for i, j in b2])]


CASES = [
    ("T2_q1", 12, 6, [("t", _translate(12, 2), 1)]),
    ("T2_q0_P_m", 12, 6, [("t", _translate(12, 2), 0), ("p", _reflect(12), 1)]),
    ("T3_pi_P_Z", 12, 6, [("t", _translate(12, 3), 2), ("p", _reflect(12), 0), ("z", _invert(12), 1)]),
    ("T4_complex", 12, 6, [("t", _translate(12, 4), 1)]),
    ("pairswap_Z", 14, 7, [("s", _pair_swap(14), 1), ("z", _invert(14), 0)]),
    ("Z_then_T2", 12, 6, [("z", _invert(12), 1), ("t", _translate(12, 2), 3)]),
    ("onebased_T1", 10, 5, [("t", [i % 10 + 1 for i in range(1, 11)], 3)]),
    ("no_nup_sublattice_flip", 10, None, [("t", _translate(10, 2), 1),
                                          ("f", [-(i + 1) if i % 2 else i for i in range(10)], 1)]),
]


def _pair(name, L, nup, blocks):
    sec = oracle.Sector.general(L, nup=nup, pauli=False, blocks=[(m, q) for _, m, q in blocks])
    b = qsgpu.SpinBasisGeneral(L, nup=nup, pauli=False, block_order=[n for n, _, _ in blocks],
                               **{n: (m, q) for n, m, q in blocks})
    return sec, b


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_general_basis_bit_exact(gpu, case):
    name, L, nup, blocks = case
    print("general basis", name, flush=True)
    sec, b = _pair(*case)
    assert b.Ns == sec.Ns
    assert np.array_equal(b.states, sec.states())
    orb, ref = b.norms()
    orb_o, ref_o = sec.norms()
    assert np.array_equal(orb, orb_o)
    assert np.array_equal(ref.view(np.uint64), ref_o.view(np.uint64))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_general_matvec_and_eigsh(gpu, case):
    name, L, nup, blocks = case
    print("general matvec", name, flush=True)
    sec, b = _pair(*case)
    terms = _terms(L)
    ref = oracle.Operator(sec, terms, check_symm=False)
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False)
    rng = np.random.default_rng(41)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    want = ref.matvec(x)
    got = op.matvec(x)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    if not b.is_complex:
        xr = x.real.copy()
        assert np.linalg.norm(op.matvec(xr) - ref.matvec(xr).real) <= 1e-12 * np.linalg.norm(want)
    # every variant of the parent-table path
    for v in (qsgpu.VARIANT_HOP, qsgpu.VARIANT_HOP + 1, qsgpu.VARIANT_HOP + 3):
        op.set_variant(v)
        assert np.linalg.norm(op.matvec(x) - want) <= 1e-12 * np.linalg.norm(want), v
    op.set_variant(0)
    H = ref.dense()
    w = np.linalg.eigvalsh(H)
    k = min(3, b.Ns - 2)
    vals = qsgpu.eigsh(op, k=k, which="SA", tol=1e-12, return_eigenvectors=False)
    assert np.allclose(vals, w[:k], atol=1e-10, rtol=0), (vals, w[:k])


def test_general_errors(gpu):
    L = 8
    with pytest.raises(qsgpu.ArgumentError, match="permutation"):
        qsgpu.SpinBasisGeneral(L, nup=4, t=([0, 0, 1, 2, 3, 4, 5, 6], 0))
    with pytest.raises(qsgpu.ArgumentError, match="particle sector"):
        qsgpu.SpinBasisGeneral(L, nup=3, z=(_invert(L), 0))
    with pytest.raises(qsgpu.ArgumentError, match="one entry per site"):
        qsgpu.SpinBasisGeneral(L, nup=4, t=([1, 2, 3], 0))
    with pytest.raises(qsgpu.ArgumentError, match="unknown symmetry block"):
        qsgpu.SpinBasisGeneral(L, nup=4, block_order=["q"], t=(_translate(L, 1), 0))
    b = qsgpu.SpinBasisGeneral(L, nup=4, t=(_translate(L, 1), 0))
    # a three-site coupling is not a two-bit hop
    with pytest.raises(qsgpu.ArgumentError, match="two-bit hop"):
        qsgpu.QuantumLinearOperator(b, [("+z-", [(0.5, 1, 2, 3)])], check_symm=False, check_herm=False)
