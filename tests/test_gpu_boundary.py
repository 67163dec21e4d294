"""GPU tests of boundary details through the C ABI.

* state_index / representative / get_amp / normalization (basis.jl:1016-1027,
  discrete_basis.jl:2386-2422) against the oracle's projector: exact column
  and label, amplitudes within 1e-13;
* the 'y' operator (operators.jl:343-379), which makes the operator complex
  (operators.jl:329-341), against dense Kronecker products;
* the check_herm / check_pcon / check_symm rejections and the real-diagonal
  rule (operators.jl:2500-2534) with the reference's messages;
* term consolidation (operators.jl:214-263): duplicates summed, zeros dropped;
* eigsh which = SM and BE (operators.jl:3367-3408) against dense spectra;
* HIP-graph replay of the halving recursion on a parent-table sector of
  2^17 or more rows, where every product runs the pre-scale pass: replays
  equal direct launches bitwise.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import _dense
import oracle
import qsgpu

pytestmark = pytest.mark.gpu


def _id(c):
    L, nup, blocks = c
    return "L%d_n%s_%s" % (L, nup, "_".join("%s%d" % (k[0], v) for k, v in blocks.items()) or "plain")


LOOKUP = [(10, 5, dict(kblock=0, pblock=1, zblock=1)), (12, 6, dict(kblock=5)),
          (12, 6, dict(kblock=6, pblock=-1, zblock=-1)), (12, 6, dict(zblock=1)), (12, 4, dict(kblock=0, pblock=1)),
          (12, 6, {}), (10, None, dict(kblock=1)), (12, 6, dict(kblock=0, a=2)), (12, 6, dict(pblock=-1))]


@pytest.mark.parametrize("case", LOOKUP, ids=_id)
def test_lookup_matches_reference_definitions(gpu, case):
    L, nup, blocks = case
    print("lookup", _id(case), flush=True)
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    labels = sec.states()
    assert b.states.tolist() == labels.tolist()
    # state_index of every label is its position
    idx, _ = b.index(labels)
    assert idx.tolist() == list(range(b.Ns))
    assert b.state_index(int(labels[-1])) == b.Ns - 1
    parent = sec.parent_states()
    if not sec.reduced:
        idx, amp = b.index(parent)
        assert idx.tolist() == list(range(b.Ns))
        assert np.all(amp == 1)
        assert b.representative(int(parent[3])) == int(parent[3])
        return
    # every parent state: its column, amplitude, representative and
    # normalization against the projector's row (one entry per row)
    colptr, row, val = sec.projector()
    col_of = np.full(sec.npar, -1, dtype=np.int64)
    amp_of = np.zeros(sec.npar, dtype=np.complex128)
    for c in range(sec.Ns):
        col_of[row[colptr[c]:colptr[c + 1]]] = c
        amp_of[row[colptr[c]:colptr[c + 1]]] = val[colptr[c]:colptr[c + 1]]
    idx, amp = b.index(parent)
    assert idx.tolist() == col_of.tolist()
    ok = col_of >= 0
    assert np.abs(amp[ok] - amp_of[ok]).max() <= 1e-13
    assert np.all(amp[~ok] == 0)
    orb, ref = b.norms()
    for s in parent[ok][:: max(1, int(ok.sum()) // 40)]:
        c = col_of[np.searchsorted(parent, s)]
        assert b.representative(int(s)) == int(labels[c])
        a = b.get_amp(int(s))
        assert abs(a - amp_of[np.searchsorted(parent, s)]) <= 1e-13
        n = b.normalization(int(s))
        if int(s) == int(labels[c]):
            assert n == ref[c]   # bit-exact at the labels
        assert abs(n - 1.0 / abs(a) ** 2) <= 1e-12 * n
    # a state whose orbit the sector drops has zero weight
    if not ok.all():
        with pytest.raises(qsgpu.ArgumentError):
            b.representative(int(parent[~ok][0]))
    with pytest.raises(qsgpu.ArgumentError):
        b.state_index(int((1 << L) - 1) if nup is not None else 1 << L)


def test_y_operator_against_kronecker_products(gpu):
    # 'y' flips its site with amplitude +i/2 from down, -i/2 from up
    # (pauli=false); a 'y' term makes the element type complex
    L = 8
    print("y operator", flush=True)
    terms = [("yy", [(0.7, i, i % L + 1) for i in range(1, L + 1)]),
             ("xy", [(0.3, i, (i + 1) % L + 1) for i in range(1, L + 1)]),
             ("zyz", [(0.2, 1, 2, 3)]), ("y", [(0.45, 5)]), ("zz", [(1.0, i, i % L + 1) for i in range(1, L + 1)])]
    for pauli in (False, True):
        b = qsgpu.SpinBasis1D(L, pauli=pauli)
        op = qsgpu.QuantumLinearOperator(b, terms, check_herm=False, check_pcon=False)
        assert np.issubdtype(op.dtype, np.complexfloating)
        H = _dense.full_hamiltonian(L, terms, pauli)
        rng = np.random.default_rng(3)
        x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
        assert np.linalg.norm(op.matvec(x) - H @ x) <= 1e-12 * np.linalg.norm(H @ x)
        ref = oracle.Operator(oracle.Sector(L, pauli=pauli), terms)
        assert np.linalg.norm(op.matvec(x) - ref.matvec(x)) <= 1e-12 * np.linalg.norm(H @ x)
    # Hermitian 'y' couplings in a symmetry sector (P'HP through the group average)
    b = qsgpu.SpinBasis1D(L, kblock=2)
    herm = [("yy", [(0.7, i, i % L + 1) for i in range(1, L + 1)]), ("z", [(0.3, 2)])]
    op = qsgpu.QuantumLinearOperator(b, herm, check_symm=False, check_pcon=False)
    M = oracle.Operator(oracle.Sector(L, kblock=2), herm, check_symm=False).dense()
    x = np.random.default_rng(4).standard_normal(b.Ns) + 0j
    assert np.linalg.norm(op.matvec(x) - M @ x) <= 1e-12 * np.linalg.norm(M @ x)


def test_check_rejections_carry_the_reference_messages(gpu):
    b = qsgpu.SpinBasis1D(8, nup=4, pauli=False, kblock=0)
    with pytest.raises(qsgpu.ArgumentError, match="operator list is not Hermitian"):
        qsgpu.QuantumLinearOperator(b, [("+-", [(0.5, 1, 2)])], check_symm=False)
    with pytest.raises(qsgpu.ArgumentError, match="violates the selected particle sector"):
        qsgpu.QuantumLinearOperator(b, [("x", [(1.0, 1)])], check_symm=False)
    with pytest.raises(qsgpu.ArgumentError, match="violates the selected symmetry sector"):
        qsgpu.QuantumLinearOperator(b, [("z", [(0.3, 1)])])
    with pytest.raises(qsgpu.ArgumentError, match="requires a real diagonal"):
        qsgpu.QuantumLinearOperator(b, [("zz", [(1.0, 1, 2)])], diagonal=np.full(b.Ns, 1j), check_symm=False)
    # with the checks off the same lists build, and act as P'HP
    op = qsgpu.QuantumLinearOperator(b, [("+-", [(0.5, 1, 2)])], check_symm=False, check_herm=False)
    assert not op.ishermitian()
    ref = oracle.Operator(oracle.Sector(8, nup=4, pauli=False, kblock=0), [("+-", [(0.5, 1, 2)])], check_symm=False)
    x = np.random.default_rng(1).standard_normal(b.Ns) + 0j
    assert np.linalg.norm(op.matvec(x) - ref.matvec(x)) <= 1e-13 * max(1.0, np.linalg.norm(ref.matvec(x)))
    op = qsgpu.QuantumLinearOperator(b, [("z", [(0.3, 1)])], check_symm=False)
    assert np.allclose(op.matvec(x), (0.3 / 8) * (4 - 4) * x, atol=1e-14)


def test_duplicate_couplings_consolidate(gpu):
    # operators.jl:214-263: (op, sites) duplicates summed in first-seen order,
    # couplings that sum to zero dropped
    L = 10
    b = qsgpu.SpinBasis1D(L, nup=5, pauli=False, kblock=0, pblock=1)
    base = _dense.xxz_terms(L, J=1.0, delta=0.6)
    dup = base + [("zz", [(0.25, 1, 2), (-0.25, 1, 2)]), ("+-", [(0.2, 3, 4)]), ("+-", [(-0.2, 3, 4)]),
                  ("zz", [(0.5, 2, 3)])]
    merged = base + [("zz", [(0.5, 2, 3)])]
    x = np.random.default_rng(8).standard_normal(b.Ns)
    a = qsgpu.QuantumLinearOperator(b, dup, check_symm=False).matvec(x)
    c = qsgpu.QuantumLinearOperator(b, merged, check_symm=False).matvec(x)
    assert np.linalg.norm(a - c) <= 1e-14 * np.linalg.norm(c)
    ref = oracle.Operator(oracle.Sector(L, nup=5, pauli=False, kblock=0, pblock=1), dup, check_symm=False)
    assert np.linalg.norm(a - ref.matvec(x).real) <= 1e-12 * np.linalg.norm(c)


def test_complex_element_type_flag(gpu):
    # complex coefficients with zero imaginary parts still make T complex
    b = qsgpu.SpinBasis1D(8, nup=4, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, [("zz", [(1.0 + 0j, i, i % 8 + 1) for i in range(1, 9)])])
    assert np.issubdtype(op.dtype, np.complexfloating)
    op = qsgpu.QuantumLinearOperator(b, [("zz", [(1.0, i, i % 8 + 1) for i in range(1, 9)])])
    assert not np.issubdtype(op.dtype, np.complexfloating)


@pytest.mark.parametrize("which", ["SM", "BE"])
def test_eigsh_sm_be_against_dense(gpu, which):
    L, nup, blocks = 12, 6, dict(kblock=0, pblock=1)
    print("eigsh", which, flush=True)
    terms = _dense.xxz_terms(L, J=1.0, delta=0.7, J2=0.2)
    D = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms).dense().real
    w = np.linalg.eigvalsh(D)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms)
    v0 = np.random.default_rng(6).standard_normal(b.Ns)
    for k in (3, 4):
        vals, vecs = qsgpu.eigsh(op, k=k, which=which, v0=v0, tol=1e-12, maxiter=3000, ncv=40)
        if which == "SM":
            want = w[np.argsort(np.abs(w), kind="stable")[:k]]
            # sorted by magnitude, as the reference re-sorts ARPACK's output
            assert np.all(np.diff(np.abs(vals)) >= -1e-12)
        else:
            hi = (k + 1) // 2
            want = np.concatenate([w[:k - hi], w[len(w) - hi:]])
        assert np.allclose(np.sort(vals), np.sort(want), atol=1e-10, rtol=0), (k, vals, want)
        for i in range(k):
            r = D @ vecs[:, i] - vals[i] * vecs[:, i]
            assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(vecs[:, i])


def test_expm_graph_replay_on_prescaled_parent_table_sector(gpu):
    # L=26 nup=13 k=0 p=+1 (about 2.6e5 rows >= 2^17): every product of the
    # captured factorization includes the pre-scale pass into the stream's
    # buffer; three replays of the captured recursion equal direct launches
    # bitwise, and both agree with the numpy restatement on a small time step
    import torch
    print("graph replay prescaled", flush=True)
    L = 26
    b = qsgpu.SpinBasis1D(L, nup=13, pauli=False, kblock=0, pblock=1)
    assert b.Ns >= 1 << 17
    terms = _dense.xxz_terms(L, J=1.0, delta=0.8, J2=0.3)
    n = b.Ns
    v = np.sin(np.arange(1, n + 1)) + 1j * np.cos(2 * np.arange(1, n + 1))
    v /= np.linalg.norm(v)
    times = [1.5, 4.0]
    direct = qsgpu.evolve(qsgpu.QuantumLinearOperator(b, terms), v, 0.0, times)
    side = torch.cuda.Stream()
    op = qsgpu.QuantumLinearOperator(b, terms).set_graphs(True)
    op.set_stream(side.cuda_stream)
    for rep in range(3):
        graphed = qsgpu.evolve(op, v, 0.0, times)
        torch.cuda.synchronize()
        assert np.array_equal(graphed, direct), rep
    # unitary evolution keeps the norm
    assert np.allclose(np.linalg.norm(direct, axis=0), 1.0, atol=1e-10)
