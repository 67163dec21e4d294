"""GPU tests of wide bases: SpinBasisGeneral beyond 32 sites and the reference's
WideSpinBasis beyond 63 (basis.jl:361-445, 530-620), states of up to 256 sites.

Each sector is built by libqsgpu (orbit minima as four-word states on the
device, labels and normalizations from the reference's stage sequence) and by
the oracle's wide restatement (tests/test_oracle_wide.py pins it).  Bars:
states, orbit sizes and normalizations bit-exact; products (plain, transposed,
conjugated, adjoint, alpha / beta) within 1e-12 of the oracle's P' H P;
eigenvalues within 1e-10 of the oracle's sector matrix and of the closed forms
the oracle tests check.  The wide form forced at 12 sites also matches the
32-bit general path.  Labels and normalizations of columns whose stage-one
orbit holds 32 or more states are parity unpinned against the reference
itself, which normalizes such orbits with BLAS nrm2 (the oracle and the
library both use generic_norm2's sequential sum).
"""
import math

import numpy as np
import pytest

import oracle
import qsgpu

pytestmark = pytest.mark.gpu


def _t(L, d=1):
    return [(i + d) % L for i in range(L)]


def _p(L):
    return [L - 1 - i for i in range(L)]


def _z(L):
    return [-(i + 1) for i in range(L)]


def _xxz(L, J=1.0, delta=0.8, J2=0.25):
    b1 = [(i, i % L + 1) for i in range(1, L + 1)]
    b2 = [(i, (i + 1) % L + 1) for i in range(1, L + 1)]
    t = [("+-", [(J / 2, i, j) for i, j in b1]), ("-+", [(J / 2, i, j) for i, j in b1]),
         ("zz", [(J * delta, i, j) for i, j in b1])]
    if J2:
        t += [("+-", [(J2 / 2, i, j) for i, j in b2]), ("-+", [(J2 / 2, i, j) for i, j in b2]),
              ("zz", [(J2 * delta, i, j) for i, j in b2])]
    return t


def _words_int(w):
    return [int(r[0]) | int(r[1]) << 64 | int(r[2]) << 128 | int(r[3]) << 192 for r in w]


CASES = [
    # name, L, nup, blocks (name, map, q), explicit
    ("L12_T_q2_forced", 12, 6, [("t", _t(12), 2)], True),
    ("L12_T2_P_forced", 12, 6, [("t", _t(12, 2), 0), ("p", _p(12), 1)], True),
    ("L12_Z_T_forced", 12, 6, [("z", _z(12), 1), ("t", _t(12), 3)], True),
    ("L40_T_q0", 40, 3, [("t", _t(40), 0)], False),
    ("L40_T_q7_complex", 40, 3, [("t", _t(40), 7)], False),
    ("L70_T_q0", 70, 2, [("t", _t(70), 0)], False),
    ("L70_T_P_even", 70, 2, [("t", _t(70), 0), ("p", _p(70), 0)], False),
    ("L100_T2_q5", 100, 2, [("t", _t(100, 2), 5)], False),
    ("L130_T_q0", 130, 2, [("t", _t(130), 0)], False),
]


def _pair(name, L, nup, blocks, explicit):
    sec = oracle.WideSector(L, nup, pauli=False, blocks=[(m, q) for _, m, q in blocks])
    b = qsgpu.SpinBasisGeneral(L, nup=nup, pauli=False, block_order=[n for n, _, _ in blocks], _explicit=explicit,
                               **{n: (m, q) for n, m, q in blocks})
    return sec, b


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_wide_basis_bit_exact(gpu, case):
    print("wide basis", case[0], flush=True)
    sec, b = _pair(*case)
    assert b.wide
    assert b.Ns == sec.Ns
    assert _words_int(b.state_words(0)) == sec.states()
    orb, ref = b.norms()
    orb_o, ref_o = sec.norms()
    assert np.array_equal(orb, orb_o)
    assert np.array_equal(ref.view(np.uint64), ref_o.view(np.uint64))
    # orbit minima ascending, each one the least state of its orbit's column
    reps = _words_int(b.state_words(1))
    assert reps == sorted(reps) and len(set(reps)) == len(reps)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_wide_products_and_eigsh(gpu, case):
    name, L = case[0], case[1]
    print("wide products", name, flush=True)
    sec, b = _pair(*case)
    terms = _xxz(L)
    M = oracle.wide_sector_matrix(sec, terms)
    op = qsgpu.QuantumLinearOperator(b, terms)
    rng = np.random.default_rng(L)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    scale = np.linalg.norm(M @ x)
    assert np.linalg.norm(op.matvec(x) - M @ x) <= 1e-12 * scale
    for view, want in ((op.T, M.T @ x), (op.conj(), M.conj() @ x), (op.H, M.conj().T @ x)):
        assert np.linalg.norm(view.matvec(x) - want) <= 1e-12 * scale
    y0 = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    out = y0.copy()
    op.mul(out, x, alpha=0.5 - 0.25j, beta=2.0)
    assert np.linalg.norm(out - (2.0 * y0 + (0.5 - 0.25j) * (M @ x))) <= 1e-12 * scale
    if not b.is_complex:
        xr = x.real.copy()
        assert np.linalg.norm(op.matvec(xr) - (M @ xr).real) <= 1e-12 * scale
        x32 = op.matvec(xr.astype(np.float32), dtype=np.float32)
        assert x32.dtype == np.float32
        assert np.linalg.norm(x32 - (M @ xr).real) <= 1e-5 * scale
    w = np.linalg.eigvalsh(M)
    k = min(3, b.Ns - 2)
    vals = qsgpu.eigsh(op, k=k, which="SA", tol=1e-12, return_eigenvectors=False)
    assert np.allclose(vals, w[:k], atol=1e-10, rtol=0), (vals, w[:k])


def test_forced_wide_form_matches_the_32_bit_path(gpu):
    # the same SpinBasisGeneral sector through both forms: states, norms and
    # products agree (the 32-bit form is pinned against the oracle separately)
    L = 12
    kw = dict(block_order=["t", "p"], t=(_t(L, 2), 0), p=(_p(L), 1))
    a = qsgpu.SpinBasisGeneral(L, nup=6, pauli=False, **kw)
    w = qsgpu.SpinBasisGeneral(L, nup=6, pauli=False, _explicit=True, **kw)
    assert not a.wide and w.wide
    assert np.array_equal(a.states, w.states)
    assert np.array_equal(a.norms()[1].view(np.uint64), w.norms()[1].view(np.uint64))
    terms = _xxz(L)
    x = np.random.default_rng(5).standard_normal(a.Ns)
    ya = qsgpu.QuantumLinearOperator(a, terms).matvec(x)
    yw = qsgpu.QuantumLinearOperator(w, terms).matvec(x)
    assert np.linalg.norm(ya - yw) <= 1e-13 * np.linalg.norm(ya)


def test_wide_monomial_couplings_beyond_hops(gpu):
    # three-site and 'y' couplings (not two-bit hops; 'xy' leaves the particle
    # sector, its images outside it drop out) on a wide basis against the
    # oracle, a user diagonal included
    L, nup = 66, 2
    t = _t(L)
    terms = [("+z-", [(0.3, i, i % L + 1, (i + 1) % L + 1) for i in range(1, L + 1)]),
             ("-z+", [(0.3, i, i % L + 1, (i + 1) % L + 1) for i in range(1, L + 1)]),
             ("xy", [(0.2, i, i % L + 1) for i in range(1, L + 1)]),
             ("yx", [(0.2, i, i % L + 1) for i in range(1, L + 1)]),
             ("zz", [(0.7, i, i % L + 1) for i in range(1, L + 1)])]
    sec = oracle.WideSector(L, nup, pauli=True, blocks=[(t, 4)])
    b = qsgpu.SpinBasisGeneral(L, nup=nup, pauli=True, t=(t, 4))
    M = oracle.wide_sector_matrix(sec, terms)
    d = np.random.default_rng(2).standard_normal(b.Ns)
    op = qsgpu.QuantumLinearOperator(b, terms, diagonal=d, check_herm=False, check_pcon=False)
    x = np.random.default_rng(3).standard_normal(b.Ns) + 0j
    want = M @ x + d * x
    assert np.linalg.norm(op.matvec(x) - want) <= 1e-12 * np.linalg.norm(want)


def test_wide_closed_forms(gpu):
    # one magnon on 70 sites: the momentum-k sector is one state of energy
    # J cos(2 pi k / L) + Delta (L - 4) / 4; two particles on an XX ring of 80
    # sites: the eigensolver's ground state against the oracle's sector matrix
    # (whose spectra the oracle tests pin to the free-fermion pair energies)
    L, J, delta = 70, 1.0, 0.6
    for k in (0, 9, 35):
        b = qsgpu.SpinBasisGeneral(L, nup=1, pauli=False, t=(_t(L), k))
        op = qsgpu.QuantumLinearOperator(b, _xxz(L, J=J, delta=delta, J2=0.0))
        assert b.Ns == 1
        e = op.matvec(np.ones(1, dtype=np.complex128))[0]
        assert abs(e - (J * math.cos(2 * math.pi * k / L) + delta * (L - 4) / 4)) <= 1e-12
    L = 80
    b = qsgpu.SpinBasisGeneral(L, nup=2, pauli=False, t=(_t(L), 0))
    sec = oracle.WideSector(L, 2, pauli=False, blocks=[(_t(L), 0)])
    op = qsgpu.QuantumLinearOperator(b, _xxz(L, J=1.0, delta=0.0, J2=0.0))
    M = oracle.wide_sector_matrix(sec, _xxz(L, J=1.0, delta=0.0, J2=0.0))
    e0 = qsgpu.eigsh(op, k=1, which="SA", tol=1e-12, return_eigenvectors=False)[0]
    assert abs(e0 - np.linalg.eigvalsh(M)[0]) <= 1e-10


@pytest.mark.parametrize("L,blocks", [(70, [("t", _t(70), 3)]), (66, [("t", _t(66), 0), ("p", _p(66), 0)]),
                                      (12, [("t", _t(12), 2)])])
def test_wide_lookup_against_projector(gpu, L, blocks):
    # state_index / representative / get_amp (basis.jl:1049-1075,
    # discrete_basis.jl:2386-2422): every parent state's column and amplitude
    # against the oracle's projector row (one entry per row), within 1e-13
    print("wide lookup", L, flush=True)
    nup = 2
    sec = oracle.WideSector(L, nup, pauli=False, blocks=[(m, q) for _, m, q in blocks])
    b = qsgpu.SpinBasisGeneral(L, nup=nup, pauli=False, block_order=[n for n, _, _ in blocks], _explicit=True,
                               **{n: (m, q) for n, m, q in blocks})
    P = sec.projector_sparse().tocsr()
    parent = sec.parent_states()
    col_of = np.full(len(parent), -1, dtype=np.int64)
    amp_of = np.zeros(len(parent), dtype=np.complex128)
    for r in range(len(parent)):
        a, b_ = P.indptr[r], P.indptr[r + 1]
        if b_ > a:
            col_of[r] = P.indices[a]
            amp_of[r] = P.data[a]
    idx, amp = b.index(parent)
    assert idx.tolist() == col_of.tolist()
    ok = col_of >= 0
    assert np.abs(amp[ok] - amp_of[ok]).max() <= 1e-13
    labels = sec.states()
    assert [b.state_index(s) for s in labels[:5]] == list(range(5))
    for r in np.flatnonzero(ok)[:: max(1, int(ok.sum()) // 20)]:
        assert b.representative(parent[r]) == labels[col_of[r]]
        assert abs(b.get_amp(parent[r]) - amp_of[r]) <= 1e-13
    # outside the particle sector, and beyond the chain
    idx, amp = b.index([1, (1 << L) | 1])
    assert idx.tolist() == [-1, -1] and np.all(amp == 0)


def test_wide_lookup_entry_on_a_64_bit_basis(gpu):
    # qs_basis_lookup_wide on a SpinBasis1D sector equals qs_basis_lookup
    import ctypes
    from qsgpu import _lib
    b = qsgpu.SpinBasis1D(12, nup=6, pauli=False, kblock=1)
    parent = oracle.fixed_weight_states(12, 6)
    idx1, amp1 = b.index(parent)
    words = np.zeros((len(parent) + 1, 4), dtype=np.uint64)
    words[:-1, 0] = parent
    words[-1, 0], words[-1, 1] = parent[0], 1          # a bit beyond 64 sites: invalid
    idx = np.empty(len(words), dtype=np.int64)
    amp = np.empty(len(words), dtype=np.complex128)
    qsgpu.errors.check(_lib.lib().qs_basis_lookup_wide(b._h, words.ctypes.data, len(words), idx.ctypes.data,
                                                        amp.ctypes.data))
    assert idx[:-1].tolist() == idx1.tolist() and idx[-1] == -1
    assert np.array_equal(amp[:-1], amp1) and amp[-1] == 0


def test_wide_errors(gpu):
    with pytest.raises(qsgpu.ArgumentError, match="require nup"):
        qsgpu.SpinBasisGeneral(70, t=(_t(70), 0))
    with pytest.raises(qsgpu.ArgumentError, match="between 1 and 256"):
        qsgpu.SpinBasisGeneral(300, nup=1, t=(_t(300), 0))
    with pytest.raises(qsgpu.ArgumentError, match="particle sector"):
        qsgpu.SpinBasisGeneral(70, nup=3, z=(_z(70), 0))
    # two particles at zero momentum are all reflection-even: the odd sector is empty
    with pytest.raises(qsgpu.ArgumentError, match="empty"):
        qsgpu.SpinBasisGeneral(70, nup=2, block_order=["t", "p"], t=(_t(70), 0), p=(_p(70), 1))
    b = qsgpu.SpinBasisGeneral(70, nup=2, t=(_t(70), 0))
    assert len(b.reps) == b.Ns                    # Python ints through the four-word export
    with pytest.raises(qsgpu.ArgumentError, match="four words"):
        qsgpu.SpinBasis1D.reps.fget(b)
    op = qsgpu.QuantumLinearOperator(b, _xxz(70))
    with pytest.raises(qsgpu.ArgumentError, match="one matvec kernel"):
        op.set_variant(qsgpu.VARIANT_HOP)
    with pytest.raises(qsgpu.ArgumentError, match="symmetry sector"):
        qsgpu.QuantumLinearOperator(b, [("zz", [(1.0, 1, 2)])])
