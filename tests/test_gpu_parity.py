"""GPU parity: libqsgpu (through its C ABI) against the CPU oracle.

Bars: integer basis work bit-exact (states, order, orbit sizes, reference
labels and normalizations); FP64 matvec within 1e-12 relative; FP32-stream
matvec within 2e-6 relative; eigenvalues within 1e-10 (FP64) of dense /
oracle results; Krylov exponentials within 3e-11 (the reference's own
evolve-vs-dense tolerance, completeness_gaps.jl:241-253).
"""
import numpy as np
import pytest

import _dense
import oracle
import qsgpu

pytestmark = pytest.mark.gpu

SECTORS = [
    (8, 4, {}), (8, None, {}), (8, 4, dict(kblock=0)), (8, 4, dict(kblock=1)), (8, 4, dict(kblock=3)),
    (8, 4, dict(kblock=4)), (8, 4, dict(pblock=1)), (8, 4, dict(pblock=-1)), (8, 4, dict(zblock=1)),
    (8, 4, dict(zblock=-1)), (8, 4, dict(kblock=0, pblock=1)), (8, 4, dict(kblock=4, pblock=-1, zblock=1)),
    (8, 4, dict(kblock=0, pblock=1, zblock=1)), (8, 4, dict(kblock=2, zblock=-1)), (8, 4, dict(pzblock=1)),
    (8, 4, dict(kblock=0, pzblock=-1)), (8, None, dict(zAblock=1)), (8, None, dict(zBblock=-1, zAblock=-1)),
    (10, 5, dict(kblock=0, pblock=1, zblock=1)), (10, 5, dict(kblock=3, zblock=1)), (12, 6, dict(kblock=0, pblock=-1, zblock=-1)),
    (12, 6, dict(kblock=6, pblock=1, zblock=1)), (12, 6, dict(kblock=5)), (12, 4, dict(kblock=0, pblock=1)),
    (12, 6, dict(zblock=1)), (12, 6, dict(kblock=0, a=2)), (10, None, dict(kblock=1)), (16, 8, {}),
    (16, 8, dict(kblock=0, pblock=1, zblock=1)), (20, 10, dict(kblock=0, pblock=1, zblock=1)),
    (18, 9, dict(kblock=3, zblock=1)),
]


def _id(c):
    L, nup, b = c
    return "L%d_n%s_%s" % (L, nup, "_".join("%s%s" % (k[0], v) for k, v in sorted(b.items())) or "plain")


def _column_minima(sec):
    colptr, row, _ = sec.projector()
    par = sec.parent_states()
    return np.array([par[row[colptr[c]:colptr[c + 1]]].min() for c in range(sec.Ns)], dtype=np.uint64)


def _terms(L, periodic=True):
    return _dense.xxz_terms(L, J=1.0, delta=0.7, J2=0.35, periodic=periodic)


@pytest.mark.parametrize("case", SECTORS, ids=_id)
def test_basis_bit_exact(gpu, case):
    L, nup, blocks = case
    print("basis", _id(case), flush=True)
    ref = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    assert b.Ns == ref.Ns
    if ref.reduced:
        assert b.reps.tolist() == _column_minima(ref).tolist()
    else:
        assert b.reps.tolist() == ref.parent_states().tolist()
    assert b.states.tolist() == ref.states().tolist()
    orb, nrm = b.norms()
    orb_ref, nrm_ref = ref.norms()
    assert orb.tolist() == orb_ref.tolist()
    assert nrm.tobytes() == nrm_ref.tobytes()


@pytest.mark.parametrize("case", SECTORS, ids=_id)
def test_matvec_fp64_matches_oracle(gpu, case):
    L, nup, blocks = case
    print("matvec", _id(case), flush=True)
    terms = _terms(L)
    if nup is None:
        terms = terms + [("x", [(0.3, i) for i in range(1, L + 1)])]
    ref_s = oracle.Sector(L, nup=nup, pauli=True, **blocks)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=True, **blocks)
    ref = oracle.Operator(ref_s, terms, check_symm=False)
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False, check_pcon=nup is not None)
    rng = np.random.default_rng(L * 100 + b.Ns)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    want = ref.matvec(x)
    got = op.matvec(x.astype(np.complex128))
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    if not b.is_complex:
        xr = x.real.copy()
        got_r = op.matvec(xr)
        assert got_r.dtype == np.float64
        assert np.linalg.norm(got_r - ref.matvec(xr).real) <= 1e-12 * np.linalg.norm(want)
        got32 = op.matvec(xr.astype(np.float32), dtype=np.float32)
        assert np.linalg.norm(got32 - ref.matvec(xr).real) <= 2e-6 * np.linalg.norm(want)
    got64c = op.matvec(x.astype(np.complex64), dtype=np.complex64)
    assert np.linalg.norm(got64c - want) <= 2e-6 * np.linalg.norm(want)


@pytest.mark.parametrize("hop", [False, True], ids=["general", "hop"])
@pytest.mark.parametrize("blocks", [dict(kblock=1), dict(kblock=0), dict(kblock=2, zblock=-1), {},
                                    dict(zblock=1)], ids=str)
def test_matvec_nonsymmetric_terms_flags_and_diagonal(gpu, blocks, hop):
    # P'HP of terms that break the symmetry (check_symm=false), a complex
    # non-Hermitian term list, a user diagonal, alpha/beta, transpose/conj/adjoint.
    # Without the three-site term every coupling is a two-bit hop, so the
    # default variant is the hop kernel.
    L, nup = 10, 5
    terms = [("+-", [(0.7 + 0.2j, 1, 2), (0.4, 3, 4)]), ("-+", [(0.3, 2, 5)]), ("z", [(0.9, 1), (-0.2, 7)]),
             ("zz", [(0.55, 2, 9)])]
    if not hop:
        terms.append(("+z-", [(0.25j, 3, 6, 8)]))
    print("flags", blocks, hop, flush=True)
    ref_s = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    rng = np.random.default_rng(11)
    diag = rng.standard_normal(b.Ns) + 0.5j * rng.standard_normal(b.Ns)
    ref = oracle.Operator(ref_s, terms, check_symm=False, diagonal=diag)
    op = qsgpu.QuantumLinearOperator(b, terms, diagonal=diag, check_symm=False, check_herm=False)
    if hop:
        op.set_variant(qsgpu.VARIANT_HOP + 1)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    variants = [None] + ([qsgpu.VARIANT_SPLIT] if hop and set(blocks) <= {"zblock"} else [])
    for v in variants:
        if v is not None:
            op.set_variant(v)
        _flags_and_rhs(op, ref, b, x, rng)


def _flags_and_rhs(op, ref, b, x, rng):
    y0 = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    for t, c in [(False, False), (True, False), (False, True), (True, True)]:
        want = ref.matvec(x, y=y0, alpha=1.25 - 0.5j, beta=-0.5 + 0.25j, transposed=t, conjugated=c)
        opf = op if not (t or c) else (op.transpose() if t and not c else (op.conj() if c and not t else op.adjoint()))
        out = y0.copy()
        opf.mul(out, x, alpha=1.25 - 0.5j, beta=-0.5 + 0.25j)
        assert np.linalg.norm(out - want) <= 1e-12 * np.linalg.norm(want), (t, c)
    # a matrix right-hand side applies column by column (operators.jl:3185-3208)
    X = np.asfortranarray(rng.standard_normal((b.Ns, 3)) + 1j * rng.standard_normal((b.Ns, 3)))
    Y = np.zeros_like(X)
    op.mul(Y, X)
    for j in range(3):
        assert np.linalg.norm(Y[:, j] - ref.matvec(X[:, j])) <= 1e-12 * np.linalg.norm(Y[:, j])


@pytest.mark.parametrize("case", [(12, 6, dict(kblock=0, pblock=1, zblock=1)), (12, 6, dict(kblock=3)),
                                  (12, 6, {}), (14, 7, dict(zblock=-1)), (12, 6, dict(kblock=6, pblock=-1))], ids=_id)
@pytest.mark.parametrize("which", ["SA", "LA", "LM"])
def test_eigsh_fp64_against_dense(gpu, case, which):
    L, nup, blocks = case
    print("eigsh", _id(case), which, flush=True)
    terms = _terms(L)
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms, check_symm=False).dense()
    ev = np.linalg.eigvalsh(ref)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms)
    k = 4
    vals, vecs = qsgpu.eigsh(op, k=k, which=which)
    if which == "SA":
        want = ev[:k]
    elif which == "LA":
        want = ev[::-1][:k]
    else:
        want = ev[np.argsort(-np.abs(ev), kind="stable")][:k]
    assert np.allclose(vals, want, atol=1e-10, rtol=0)
    for j in range(k):
        r = ref @ vecs[:, j] - vals[j] * vecs[:, j]
        assert np.linalg.norm(r) < 1e-8


def test_eigsh_c1_against_arpack(gpu):
    # L=16 nup=8 XXZ Delta=0.5, eigsh k=4 :SA against scipy's ARPACK on the oracle CSR
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    L = 16
    terms = [("+-", [(0.5, i, i % L + 1) for i in range(1, L + 1)]), ("-+", [(0.5, i, i % L + 1) for i in range(1, L + 1)]),
             ("zz", [(0.5, i, i % L + 1) for i in range(1, L + 1)])]
    sec = oracle.Sector(L, nup=8, pauli=False)
    rowptr, col, val = oracle.CSR(sec, terms).export()
    A = sp.csr_matrix((val, col, rowptr), shape=(sec.Ns, sec.Ns))
    want = np.sort(sla.eigsh(A, k=4, which="SA", tol=0)[0])
    b = qsgpu.SpinBasis1D(L, nup=8, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, terms)
    rng = np.random.default_rng(0x51a7)
    v0 = rng.standard_normal(b.Ns)
    vals, vecs = qsgpu.eigsh(op, k=4, which="SA", v0=v0 / np.linalg.norm(v0))
    assert np.allclose(vals, want, atol=1e-10, rtol=0)
    x = rng.standard_normal(b.Ns)
    assert np.linalg.norm(op.matvec(x) - A @ x) <= 1e-12 * np.linalg.norm(A @ x)


@pytest.mark.parametrize("case", [(10, 5, dict(kblock=0, pblock=1, zblock=1)), (10, 5, dict(kblock=2)), (8, 4, {})],
                         ids=_id)
def test_expm_against_dense(gpu, case):
    import scipy.linalg as sl
    L, nup, blocks = case
    print("expm", _id(case), flush=True)
    terms = _dense.xxz_terms(L, J=1.0, delta=0.9)
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms).dense()
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms)
    n = b.Ns
    v = np.sin(np.arange(1, n + 1)) + 1j * np.cos(2 * np.arange(1, n + 1))
    v /= np.linalg.norm(v)
    times = [0.0, 0.17, 0.83, 3.0]
    out = qsgpu.evolve(op, v, 0.0, times)
    for j, t in enumerate(times):
        want = sl.expm(-1j * t * ref) @ v
        assert np.linalg.norm(out[:, j] - want) < 3e-11
        assert abs(np.linalg.norm(out[:, j]) - 1) < 3e-12
    e = qsgpu.ExpmMultiplyParallel(op, -0.37j)
    assert np.linalg.norm(e.apply(v) - sl.expm(-0.37j * ref) @ v) < 3e-11
    # imaginary time renormalises (operators.jl:2157-2161)
    out = qsgpu.evolve(op, v, 0.0, [0.5], imag_time=True)
    want = sl.expm(-0.5 * ref) @ v
    assert np.linalg.norm(out[:, 0] - want / np.linalg.norm(want)) < 3e-11


def test_fullsize_c2_basis_and_ground_state(gpu):
    # J1-J2 XXZ L=28, nup=14, k=0, p=+1, z=+1 (361,270 rows).  Size-independent properties:
    # dimension from the character formula, orbit sizes summing to C(28,14),
    # Hermiticity <x,Hy> = <Hx,y>, an FP64 ground state with residual
    # <= 1e-8 |E0| and a Temple-bound error <= 1e-10, and the FP32-stream
    # eigenvalue within 1e-8 of the FP64 one.  The basis arrays themselves are
    # checked against the oracle's digests in test_fullsize_basis_digests.
    import math
    L = 28
    print("c2 basis", flush=True)
    b = qsgpu.SpinBasis1D(L, nup=14, pauli=False, kblock=0, pblock=1, zblock=1)
    assert b.Ns == 361270
    orb, _ = b.norms()
    assert int(orb.astype(np.int64).sum()) == math.comb(28, 14)
    reps = b.reps
    assert np.all(np.diff(reps.astype(np.int64)) > 0)
    terms = _dense.xxz_terms(L, J=1.0, delta=1.0, J2=0.4)
    op = qsgpu.QuantumLinearOperator(b, terms)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(b.Ns), rng.standard_normal(b.Ns)
    assert abs(x @ op.matvec(y) - op.matvec(x) @ y) < 1e-10 * abs(x @ op.matvec(y)) + 1e-9
    # at this size the matrix-free product gathers a pre-scaled x; the
    # assembled rows carry the unscaled weights of the pull enumeration
    print("c2 assembled vs matrix-free", flush=True)
    op_a = qsgpu.QuantumLinearOperator(b, terms).assemble()
    op_a.set_variant(qsgpu.VARIANT_SELL)
    hy = op.matvec(y)
    assert np.linalg.norm(op_a.matvec(y) - hy) <= 1e-13 * np.linalg.norm(hy)
    del op_a
    print("c2 eigsh fp64", flush=True)
    v0 = rng.standard_normal(b.Ns)
    e64, v64 = qsgpu.eigsh(op, k=2, which="SA", tol=1e-10, v0=v0)
    # residuals of the unit Ritz vectors, with |E0| <= ||H||
    res = [np.linalg.norm(op.matvec(v64[:, i]) - e64[i] * v64[:, i]) / np.linalg.norm(v64[:, i]) for i in (0, 1)]
    assert res[0] <= 1e-8 * abs(e64[0]), res
    # Temple's inequality: with the second eigenvalue at least E1 - r1 (the
    # interval [E1 - r1, E1 + r1] holds an eigenvalue and the k=2 solve found
    # the two lowest), the ground-state energy lies in
    # [E0 - r0^2 / (E1 - r1 - E0), E0]
    gap = e64[1] - res[1] - e64[0]
    assert gap > 0, (e64, res)
    assert res[0] ** 2 / gap <= 1e-10, (res, gap)
    print("c2 eigsh fp32", flush=True)
    e32, _ = qsgpu.eigsh(op, k=1, which="SA", tol=1e-10, v0=v0, dtype=np.float32)
    assert abs(e32[0] - e64[0]) <= 1e-8


def _digest(a, dtype):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.dtype(dtype).newbyteorder("<")).tobytes()).hexdigest()


def _digest_cases():
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "sector_digests.json")
    return json.load(open(path))["sectors"]


@pytest.mark.parametrize("rec", _digest_cases(), ids=lambda r: "L%d_%s" % (r["L"], "_".join(
    "%s%d" % (k[0], v) for k, v in r["blocks"].items())))
def test_fullsize_basis_digests(gpu, rec):
    # bit-exact basis at full size: states(basis) (the reference labels), the
    # orbit sizes and normalization(basis, label) of the L=28 ground-state
    # sector and of the 26 L=24 quench blocks, against SHA-256 digests of the
    # oracle's arrays (tests/golden/make_digests.py)
    print("digests", rec["L"], rec["blocks"], flush=True)
    b = qsgpu.SpinBasis1D(rec["L"], nup=rec["nup"], pauli=False, **rec["blocks"])
    assert b.Ns == rec["Ns"]
    states = b.states
    assert [int(s) for s in states[:4]] == rec["first_states"]
    assert _digest(states, np.uint64) == rec["states"]
    orb, ref = b.norms()
    assert int(orb.astype(np.int64).sum()) == rec["orbit_sum"]
    assert _digest(orb, np.uint32) == rec["orbit_size"]
    assert _digest(ref, np.float64) == rec["ref_norm"]


def test_prescaled_gather_complex_sector_fullsize(gpu):
    # L=26 nup=13 k=1 (about 4.0e5 rows, complex characters): the matrix-free
    # product gathers a pre-scaled x through the character-tracking kernel;
    # the assembled rows carry the unscaled weights.  ComplexF64 and the
    # ComplexF32 stream (FP64 pre-scaled copy) against them.
    print("prescaled complex sector", flush=True)
    L = 26
    b = qsgpu.SpinBasis1D(L, nup=13, pauli=False, kblock=1)
    assert b.Ns >= 1 << 17
    terms = _dense.xxz_terms(L, J=1.0, delta=0.7, J2=0.3)
    op = qsgpu.QuantumLinearOperator(b, terms)
    op_a = qsgpu.QuantumLinearOperator(b, terms).assemble()
    op_a.set_variant(qsgpu.VARIANT_SELL)
    rng = np.random.default_rng(13)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    want = op_a.matvec(x)
    got = op.matvec(x)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    got32 = op.matvec(x.astype(np.complex64))
    assert np.linalg.norm(got32 - want) <= 1e-6 * np.linalg.norm(want)
    # Hermitian: <x, Hx> is real
    assert abs(np.vdot(x, got).imag) <= 1e-10 * abs(np.vdot(x, got))


@pytest.mark.parametrize("case", [(12, 6, {}), (12, None, {}), (12, 6, dict(zblock=-1)), (12, None, dict(zblock=1)),
                                  (12, 6, dict(kblock=0, pblock=1, zblock=1)), (12, 6, dict(kblock=5)),
                                  (12, 6, dict(kblock=6, pblock=-1, zblock=-1))], ids=_id)
def test_every_kernel_variant_matches_oracle(gpu, case):
    # lanes-per-row variants and the autotuner change only the summation grouping;
    # long-range hops exercise the incremental combinadic rank with particles in between
    L, nup, blocks = case
    print("variants", _id(case), flush=True)
    terms = _terms(L) + [("+-", [(0.3, i, (i + 4) % L + 1) for i in range(1, L + 1)]),
                         ("-+", [(0.3, i, (i + 4) % L + 1) for i in range(1, L + 1)])]
    if nup is None:
        # pair creation / annihilation: the hop tables' equal-bit patterns
        terms = terms + [("++", [(0.2, i, i % L + 1) for i in range(1, L + 1)]),
                         ("--", [(0.2, i, i % L + 1) for i in range(1, L + 1)])]
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=True, **blocks), terms, check_symm=False)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=True, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False, check_pcon=False)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    want = ref.matvec(x)
    hops = (qsgpu.VARIANT_HOP, qsgpu.VARIANT_HOP + 1, qsgpu.VARIANT_HOP + 2, qsgpu.VARIANT_HOP + 3)
    for lpr in (1, 2, 4, 8) + hops:
        op.set_variant(lpr)
        got = op.matvec(x)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), lpr
        if not b.is_complex:
            got_r = op.matvec(x.real.copy())
            assert np.linalg.norm(got_r - ref.matvec(x.real.copy()).real) <= 1e-12 * np.linalg.norm(want), lpr
            got32 = op.matvec(x.real.astype(np.float32), dtype=np.float32)
            assert np.linalg.norm(got32 - ref.matvec(x.real.copy()).real) <= 2e-6 * np.linalg.norm(want), lpr
    split = nup is not None and set(blocks) <= {"zblock"}
    if split:
        # the split kernel (high word / low word blocks) on the same operator
        op.set_variant(qsgpu.VARIANT_SPLIT)
        got = op.matvec(x)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), "split"
        got32 = op.matvec(x.real.astype(np.float32), dtype=np.float32)
        assert np.linalg.norm(got32 - ref.matvec(x.real.copy()).real) <= 2e-6 * np.linalg.norm(want), "split"
        got64c = op.matvec(x.astype(np.complex64), dtype=np.complex64)
        assert np.linalg.norm(got64c - want) <= 2e-6 * np.linalg.norm(want), "split"
    else:
        with pytest.raises(qsgpu.ArgumentError, match="split"):
            op.set_variant(qsgpu.VARIANT_SPLIT)
    lpr, ms = op.tune(np.complex128, repeats=2)
    assert lpr in (1, 2, 4, 8) + hops + (tuple(qsgpu.VARIANT_SPLIT + k for k in range(8)) if split else ()) and ms > 0
    got = op.matvec(x)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


def test_hop_variant_rejected_for_other_couplings(gpu):
    # a three-site coupling is not a two-bit hop: the hop kernel refuses it
    b = qsgpu.SpinBasis1D(8, nup=4, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, [("+z-", [(0.5, 1, 2, 3)])], check_symm=False, check_herm=False)
    with pytest.raises(qsgpu.ArgumentError, match="hop"):
        op.set_variant(qsgpu.VARIANT_HOP)


@pytest.mark.parametrize("case", [(24, 12, dict(zblock=1)), (24, 10, {}), (20, 10, dict(zblock=-1)), (16, 7, {})],
                         ids=_id)
def test_split_kernel_against_oracle(gpu, case):
    # split point 10 for L >= 20 (rows of one block share bits 10..L-1), L/2
    # below; next-nearest and fourth-neighbour bonds cross the split, and
    # under spin inversion the bonds onto bit L-1 read inverted blocks.  Every
    # variant: the first kernel (0 or 1 workgroups per CU held back), the lean
    # kernel in one pass and its tiled two-pass form at the four tile budgets
    # (at L = 24 the smallest tiles cover 2^9 of the 2^14 high words)
    L, nup, blocks = case
    print("split", _id(case), flush=True)
    terms = _terms(L) + [("+-", [(0.3, i, (i + 3) % L + 1) for i in range(1, L + 1)]),
                         ("-+", [(0.3, i, (i + 3) % L + 1) for i in range(1, L + 1)])]
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms)
    rng = np.random.default_rng(17)
    x = rng.standard_normal(b.Ns)
    want = ref.matvec(x).real
    y0 = rng.standard_normal(b.Ns)
    for k in (0, 1, 3, 4, 5, 6, 7):
        op.set_variant(qsgpu.VARIANT_SPLIT + k)
        for dt, tol in ((np.float64, 1e-12), (np.float32, 2e-6)):
            got = op.matvec(x.astype(dt), dtype=dt)
            assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want), (k, dt)
        out = y0.copy()
        op.mul(out, x, alpha=0.75, beta=-1.5)
        assert np.linalg.norm(out - (0.75 * want - 1.5 * y0)) <= 1e-12 * np.linalg.norm(out), k


@pytest.mark.parametrize("case", [(24, 12, dict(zblock=-1)), (24, 11, {})], ids=_id)
def test_split_tiled_complex_couplings_and_flags(gpu, case):
    # complex hop amplitudes (the adjoint list differs), transposed, conjugated
    # and adjoint products and a user diagonal through both passes
    L, nup, blocks = case
    print("split tiled flags", _id(case), flush=True)
    terms = [("+-", [(0.5 + 0.2j, i, i % L + 1) for i in range(1, L + 1)]),
             ("-+", [(0.5 - 0.2j, i, i % L + 1) for i in range(1, L + 1)]),
             ("zz", [(0.8, i, i % L + 1) for i in range(1, L + 1)]),
             ("+-", [(0.1j, i, (i + 1) % L + 1) for i in range(1, L + 1)])]
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    rng = np.random.default_rng(23)
    diag = rng.standard_normal(sec.Ns)
    # complex amplitudes are not inversion-symmetric: P'HP keeps their
    # inversion average (check_symm off, as in the reference)
    ref = oracle.Operator(sec, terms, diagonal=diag, check_symm=False)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms, diagonal=diag, check_herm=False, check_symm=False)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    flags = ((False, False, op), (True, False, op.T), (False, True, op.conj()), (True, True, op.H))
    wants = [ref.matvec(x, transposed=tr, conjugated=cj) for tr, cj, _ in flags]
    for k in (0, 3, 4, 6):
        op.set_variant(qsgpu.VARIANT_SPLIT + k)
        for (tr, cj, view), want in zip(flags, wants):
            got = view.matvec(x)
            assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), (k, tr, cj)
            got64 = view.matvec(x.astype(np.complex64), dtype=np.complex64)
            assert np.linalg.norm(got64 - want) <= 2e-6 * np.linalg.norm(want), (k, tr, cj)


def test_split_kernel_fullsize_c3(gpu):
    # L=32, nup=16, z=+1 (N = 300,540,195): the split kernel equals the hop
    # kernel (validated against the oracle at small L) to 1e-12 in FP64, and
    # the FP32 stream is within 2e-6; the random fields cancel in z+.
    import torch
    print("split fullsize L=32 z+", flush=True)
    b = qsgpu.SpinBasis1D(32, nup=16, pauli=False, zblock=1)
    h = np.random.default_rng(0xC3).uniform(-2.0, 2.0, 32)
    terms = _dense.xxz_terms(32, J=1.0, delta=1.0, J2=0.0, periodic=True) + \
        [("z", [(float(h[i]), i + 1) for i in range(32)])]
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False)
    assert b.Ns == 300540195
    g = torch.Generator(device="cpu").manual_seed(7)
    x = torch.randn(b.Ns, generator=g, dtype=torch.float64).cuda()
    y_hop = torch.empty_like(x)
    y_split = torch.empty_like(x)
    op.set_variant(qsgpu.VARIANT_HOP)
    op.matvec_device(x, y_hop)
    torch.cuda.synchronize()
    ref = torch.linalg.vector_norm(y_hop).item()
    x32 = x.float()
    y32 = torch.empty_like(x32)
    for k in (0, 3, 4, 5, 6, 7):
        print("  split variant", k, flush=True)
        op.set_variant(qsgpu.VARIANT_SPLIT + k)
        op.matvec_device(x, y_split)
        torch.cuda.synchronize()
        assert torch.linalg.vector_norm(y_split - y_hop).item() <= 1e-12 * ref, k
        op.matvec_device(x32, y32)
        torch.cuda.synchronize()
        assert torch.linalg.vector_norm(y32.double() - y_hop).item() <= 2e-6 * ref, k
    # <x, Hx> is real and the product is symmetric: <x, H x32> == <x32, H x> in FP64
    del y32, x32


def test_c3_eigsh_fp32_stream_against_fp64(gpu):
    # The precision-form sector: L=32 nup=16 z=+1 (N = 300,540,195), k=4 lowest
    # eigenvalues (the random fields cancel in z+).  FP32-stored Krylov basis
    # with FP64 dots / norms / tridiagonal solve against the FP64 run: every
    # eigenvalue within 1e-8, every FP64 Ritz vector's residual within
    # 1e-8 ||H|| (|E_i| <= ||H||).
    import torch
    print("c3 eigsh", flush=True)
    b = qsgpu.SpinBasis1D(32, nup=16, pauli=False, zblock=1)
    h = np.random.default_rng(0xC3).uniform(-2.0, 2.0, 32)
    terms = _dense.xxz_terms(32, J=1.0, delta=1.0, J2=0.0, periodic=True) + \
        [("z", [(float(h[i]), i + 1) for i in range(32)])]
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False)
    v0 = np.random.default_rng(0x5EED).standard_normal(b.Ns)
    print("  fp64", flush=True)
    e64, vec = qsgpu.eigsh(op, k=4, which="SA", tol=1e-10, v0=v0)
    print("  fp64 E", e64.tolist(), qsgpu.eigsh.last_n_matvec, flush=True)
    hmax = max(abs(e64))
    y = torch.empty(b.Ns, dtype=torch.float64, device="cuda")
    for i in range(4):
        v = torch.from_numpy(np.ascontiguousarray(vec[:, i])).cuda()
        op.matvec_device(v, y)
        r = (torch.linalg.vector_norm(y - e64[i] * v) / torch.linalg.vector_norm(v)).item()
        print("  residual", i, r, flush=True)
        assert r <= 1e-8 * hmax, (i, r)
        del v
    del vec, y
    torch.cuda.empty_cache()
    print("  fp32", flush=True)
    e32 = qsgpu.eigsh(op, k=4, which="SA", tol=1e-10, v0=v0, dtype=np.float32, return_eigenvectors=False)
    print("  fp32 E", e32.tolist(), qsgpu.eigsh.last_n_matvec, flush=True)
    assert np.all(np.abs(e32 - e64) <= 1e-8), (e32, e64)
    assert np.all(np.diff(e64) >= 0)


@pytest.mark.parametrize("case", [(8, 4, {}), (8, None, {}), (10, 5, dict(kblock=0, pblock=1, zblock=1)),
                                  (10, 5, dict(kblock=3, zblock=1)), (12, 6, dict(kblock=6, pblock=-1, zblock=-1)),
                                  (12, 6, dict(zblock=1)), (12, 4, dict(kblock=0, pblock=1)),
                                  (16, 8, dict(kblock=0, pblock=1, zblock=1)), (12, 6, dict(kblock=0, a=2))],
                         ids=_id)
def test_assembled_matrix_matches_oracle(gpu, case):
    # the assembled sector matrix (operators.jl:461-576) equals P'HP entry for
    # entry, and its plain products equal the oracle's; transposed / adjoint
    # products stay matrix-free and must still agree
    L, nup, blocks = case
    print("assembled", _id(case), flush=True)
    terms = _terms(L)
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms)
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms).assemble()
    M = op.tocsr()
    D = ref.dense()
    assert np.abs(M.toarray() - D).max() <= 1e-13 * max(1.0, np.abs(D).max())
    rng = np.random.default_rng(23)
    x = rng.standard_normal(b.Ns) + (1j * rng.standard_normal(b.Ns) if b.is_complex else 0)
    want = ref.matvec(x)
    if not b.is_complex:
        want = want.real
    for v in (0, qsgpu.VARIANT_SELL):
        op.set_variant(v)
        got = op.matvec(x)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), v
        y0 = rng.standard_normal(b.Ns)
        out = y0.astype(got.dtype)
        op.mul(out, x, alpha=0.5, beta=2.0)
        assert np.linalg.norm(out - (0.5 * want + 2.0 * y0)) <= 1e-12 * np.linalg.norm(out), v
    xf = x.astype(np.complex64 if b.is_complex else np.float32)
    got32 = op.matvec(xf, dtype=xf.dtype)
    assert np.linalg.norm(got32 - want) <= 2e-6 * np.linalg.norm(want)
    adj = op.adjoint().matvec(x)
    assert np.linalg.norm(adj - D.conj().T @ x) <= 1e-12 * np.linalg.norm(adj)
    # selecting the assembled variant defers the assembly to the first plain product
    lazy = qsgpu.QuantumLinearOperator(b, terms).set_variant(qsgpu.VARIANT_SELL)
    assert np.linalg.norm(lazy.matvec(x) - want) <= 1e-12 * np.linalg.norm(want)
    assert (lazy.tocsr() != M).nnz == 0


def test_assembled_csr_c1_equals_reference_csr(gpu):
    # XXZ L=16, nup=8, pauli=false, J=1, Delta=0.5 -- the reference's CSR
    # (zero diagonals dropped) has 117,794 stored entries; the device assembly
    # exported as CSR has the same structure and values
    L = 16
    bonds = [(i, i % L + 1) for i in range(1, L + 1)]
    terms = [("+-", [(0.5, i, j) for i, j in bonds]), ("-+", [(0.5, i, j) for i, j in bonds]),
             ("zz", [(0.5, i, j) for i, j in bonds])]
    sec = oracle.Sector(L, nup=8, pauli=False)
    rp, col, val = oracle.CSR(sec, terms).export()
    b = qsgpu.SpinBasis1D(L, nup=8, pauli=False)
    M = qsgpu.QuantumLinearOperator(b, terms).assemble().tocsr()
    assert M.nnz == 117794 == len(col)
    assert np.array_equal(M.indptr, rp) and np.array_equal(M.indices, col)
    assert np.abs(M.data - val).max() <= 1e-15


def test_expm_many_times_on_matrix_cores(gpu):
    # 40 time points from one factorization: the recombination beta V c(t) is a
    # GEMM (basis x time grid) that runs on the FP64 matrix cores
    import scipy.linalg as sl
    L, nup, blocks = 14, 7, dict(kblock=0, pblock=1, zblock=1)
    print("expm 40 times", flush=True)
    terms = _dense.xxz_terms(L, J=1.0, delta=0.9)
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms).dense()
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms)
    n = b.Ns
    v = np.sin(np.arange(1, n + 1)) + 1j * np.cos(2 * np.arange(1, n + 1))
    v /= np.linalg.norm(v)
    times = np.linspace(0.0, 0.6, 40)
    assert n >= 16   # the matrix-core path needs at least one 16-row tile
    out = qsgpu.evolve(op, v, 0.0, times)
    for j, t in enumerate(times):
        want = sl.expm(-1j * t * ref) @ v
        assert np.linalg.norm(out[:, j] - want) < 3e-11, t


def test_eigsh_sixteen_vectors(gpu):
    # k = 16 Ritz vectors (and the thick restart's kept vectors) combine on the
    # FP64 matrix cores
    L, nup, blocks = 12, 6, dict(zblock=1)
    print("eigsh k=16", flush=True)
    terms = _terms(L)
    D = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms).dense().real
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, terms)
    v0 = np.random.default_rng(5).standard_normal(b.Ns)
    vals, vecs = qsgpu.eigsh(op, k=16, which="SA", v0=v0, ncv=40, tol=1e-12)
    want = np.linalg.eigvalsh(D)[:16]
    assert np.allclose(vals, want, atol=1e-10, rtol=0)
    for i in range(16):
        r = D @ vecs[:, i] - vals[i] * vecs[:, i]
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(vecs[:, i])


def test_expm_graph_replays_equal_direct_launches(gpu):
    # With graphs enabled, on a non-blocking stream, the halving recursion
    # captures its factorization once in a HIP graph and replays it for every
    # later call; each replay must give exactly what direct launches give, and
    # both match dense exponentials.
    import scipy.linalg as sl
    import torch
    L, nup, blocks = 10, 5, dict(kblock=0, pblock=1, zblock=1)
    print("expm graph replays", flush=True)
    terms = _dense.xxz_terms(L, J=1.0, delta=0.9, J2=0.3)
    ref = oracle.Operator(oracle.Sector(L, nup=nup, pauli=False, **blocks), terms).dense()
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    n = b.Ns
    v = np.sin(np.arange(1, n + 1)) + 1j * np.cos(2 * np.arange(1, n + 1))
    v /= np.linalg.norm(v)
    times = [2.0, 6.0, 9.0]
    direct = qsgpu.evolve(qsgpu.QuantumLinearOperator(b, terms), v, 0.0, times)
    side = torch.cuda.Stream()
    op = qsgpu.QuantumLinearOperator(b, terms).set_graphs(True)
    op.set_stream(side.cuda_stream)
    for rep in range(3):
        graphed = qsgpu.evolve(op, v, 0.0, times)
        torch.cuda.synchronize()
        assert np.array_equal(graphed, direct), rep
    for j, t in enumerate(times):
        want = sl.expm(-1j * t * ref) @ v
        assert np.linalg.norm(direct[:, j] - want) < 3e-11, t
