"""Pins the CPU oracle (oracle/qs_oracle.c) before it is trusted as the checker.

Each check is chosen so a plausible restatement mistake fails it: the
reference's own known answers, closed forms, and independent dense
constructions (tests/_dense.py) that share no code or algorithm with it.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle
import _dense

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


def _terms(raw):
    return [(op, [tuple(c) for c in cs]) for op, cs in raw]


@pytest.mark.parametrize("L", range(1, 13))
def test_fixed_weight_states_against_brute_force(L):
    # basis.jl:229-250: ascending popcount-w words, C(L, w) of them
    for w in range(0, L + 1):
        got = oracle.fixed_weight_states(L, w)
        want = [s for s in range(2 ** L) if bin(s).count("1") == w]
        assert got.tolist() == want
        assert len(got) == math.comb(L, w)


def test_spin_basis_known_states():
    g = GOLD["spin_basis_L4_nup2_states"]
    s = oracle.Sector(4, nup=2, pauli=False)
    assert s.states().tolist() == g["states"]


def test_xxz_L4_known_spectrum():
    g = GOLD["xxz_L4_nup2"]
    s = oracle.Sector(g["L"], nup=g["nup"], pauli=g["pauli"])
    M = oracle.Operator(s, _terms(g["terms"])).dense()
    assert np.allclose(M, M.conj().T, atol=0)
    ev = np.linalg.eigvalsh(M)
    assert np.allclose(ev, g["eigenvalues"], atol=g["atol"], rtol=g["atol"])
    assert abs(np.trace(M).real - g["trace"]) < g["atol"]
    assert abs(np.linalg.norm(M) - g["frobenius"]) < g["atol"]


def test_qlo_protocol_known_answers():
    g = GOLD["qlo_protocol_L3_nup1"]
    s = oracle.Sector(g["L"], nup=g["nup"], pauli=g["pauli"])
    terms = _terms(g["terms"])
    op = oracle.Operator(s, terms, diagonal=g["diagonal"])
    states = s.states()
    Hf = _dense.full_hamiltonian(g["L"], terms, g["pauli"])
    ref = Hf[np.ix_(states.astype(int), states.astype(int))] + np.diag(g["diagonal"])
    x = np.array(g["vector_re"]) + 1j * np.array(g["vector_im"])
    x = x / np.linalg.norm(x)
    assert np.allclose(op.matvec(x), ref @ x, atol=g["atol"])
    dest = np.array(g["destination_re"]) + 1j * np.array(g["destination_im"])
    got = op.matvec(x, y=dest, alpha=g["alpha"], beta=g["beta"])
    assert np.allclose(got, g["alpha"] * (ref @ x) + g["beta"] * dest, atol=g["atol"])
    assert np.allclose(op.matvec(x, transposed=True), ref.T @ x, atol=g["atol"])
    assert np.allclose(op.matvec(x, conjugated=True), ref.conj() @ x, atol=g["atol"])
    assert np.allclose(op.matvec(x, transposed=True, conjugated=True), ref.conj().T @ x, atol=g["atol"])


@pytest.mark.parametrize("L,nup", [(4, 2), (6, 3), (8, 4), (6, 2)])
def test_sector_spectra_union_equals_full(L, nup):
    # completeness_gaps.jl:77-111
    terms = _dense.xxz_terms(L, J=1.0, delta=1.0)
    full = np.sort(np.linalg.eigvalsh(oracle.Operator(oracle.Sector(L, nup=nup, pauli=False), terms).dense()))
    for name, vals in (("kblock", range(L)), ("pblock", (-1, 1)), ("zblock", (-1, 1))):
        if name == "zblock" and 2 * nup != L:
            continue
        spec = []
        for v in vals:
            try:
                sec = oracle.Sector(L, nup=nup, pauli=False, **{name: v})
            except ValueError:
                continue
            spec += list(np.linalg.eigvalsh(oracle.Operator(sec, terms).dense()))
        assert np.allclose(np.sort(spec), full, atol=2e-12)


@pytest.mark.parametrize("blocks", [dict(kblock=0), dict(kblock=1), dict(kblock=3), dict(pblock=1),
                                    dict(zblock=-1), dict(kblock=0, pblock=1, zblock=1),
                                    dict(kblock=4, pblock=-1, zblock=-1), dict(kblock=2, zblock=1)])
def test_projector_matches_group_construction(blocks):
    # P'P = I (completeness_gaps.jl:113-116) and the exact-arithmetic column
    # P_chi|s0>/norm on ascending orbit minima (independent construction)
    L, nup = 8, 4
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    P = sec.projector_dense()
    assert np.allclose(P.conj().T @ P, np.eye(sec.Ns), atol=3e-13)
    _, Pd, mins = _dense.sector_projector(L, nup, **blocks)
    assert sec.Ns == Pd.shape[1] == _dense.burnside_count(L, nup, **blocks)
    assert np.allclose(P, Pd, atol=1e-13)


def test_qlo_matches_independent_sector_hamiltonian():
    L, nup = 8, 4
    terms = _dense.xxz_terms(L, J=1.0, delta=0.7, J2=0.3)
    for blocks in [dict(kblock=0, pblock=1, zblock=1), dict(kblock=2), dict(kblock=4, pblock=-1)]:
        sec = oracle.Sector(L, nup=nup, pauli=True, **blocks)
        M = oracle.Operator(sec, terms, check_symm=False).dense()
        Md, _, _ = _dense.sector_hamiltonian(L, nup, terms, True, **blocks)
        assert np.allclose(M, Md, atol=1e-13)


def test_unchecked_field_in_momentum_sector_closed_form():
    # operators_algorithm_regressions.jl:106-126: a single-site z term in a
    # k=0 sector acts as its translation average (0.7/L) sum_i z_i, which is
    # (0.7/L)(nup - L/2) times the identity.
    for L, nup in [(4, 2), (4, 1), (6, 1), (8, 3)]:
        sec = oracle.Sector(L, nup=nup, pauli=False, kblock=0)
        M = oracle.Operator(sec, [("z", [(0.7, 1)])], check_symm=False).dense()
        assert np.allclose(M, (0.7 / L) * (nup - L / 2) * np.eye(sec.Ns), atol=2e-14)


def test_labels_where_ties_are_exact():
    # single-stage k=0 and z-only sectors have exactly tied |P| entries: the
    # reference label is the first row, i.e. the orbit minimum
    L = 10
    sec = oracle.Sector(L, nup=5, pauli=False, kblock=0)
    _, _, mins = _dense.sector_projector(L, 5, kblock=0)
    assert sec.states().tolist() == mins.tolist()
    sec = oracle.Sector(L, nup=5, pauli=False, zblock=-1)
    states = sec.states()
    assert all(int(s) >> (L - 1) == 0 for s in states)
    orb, ref = sec.norms()
    assert set(orb.tolist()) == {2}
    assert np.allclose(ref, 2.0, rtol=1e-15)


def test_c1_csr_structure_and_product():
    # XXZ L=16, nup=8, pauli=false, J=1, Delta=0.5; nnz = 109,824 off-diagonal + 7,970 nonzero diagonal
    L = 16
    sec = oracle.Sector(L, nup=8, pauli=False)
    terms = [("+-", [(0.5, i, i % L + 1) for i in range(1, L + 1)]), ("-+", [(0.5, i, i % L + 1) for i in range(1, L + 1)]),
             ("zz", [(0.5, i, i % L + 1) for i in range(1, L + 1)])]
    csr = oracle.CSR(sec, terms)
    assert csr.nnz == 117794
    rng = np.random.default_rng(7)
    x = rng.standard_normal(sec.Ns)
    y1 = csr.matvec(x)
    y2 = oracle.Operator(sec, terms).matvec(x).real
    assert np.allclose(y1, y2, atol=1e-12)
    assert np.allclose(csr.matvec(x, nthreads=4), y1, atol=0)


def test_burnside_dimensions_match_benchmark_sectors():
    # sector sizes of the benchmark configurations, from the character formula
    assert _dense.burnside_count(12, 6, kblock=0, pblock=1, zblock=1) == \
        oracle.Sector(12, nup=6, kblock=0, pblock=1, zblock=1).Ns
    assert math.comb(16, 8) == 12870


def test_cpu_baseline_loops_run_on_one_and_several_threads():
    # the bench's CPU baseline (the parent-space apply loop of
    # operators.jl:2554-2589) on one thread and on several: both time a
    # bounded sample and return a positive duration
    terms = _dense.xxz_terms(12, J=1.0, delta=1.0, J2=0.4)
    t1 = oracle.bench_parent_apply(12, 6, False, terms, 500, 1)
    t4 = oracle.bench_parent_apply_mt(12, 6, False, terms, 500, 4)
    assert t1 > 0 and t4 > 0


def test_whole_product_baseline_builds_the_reference_sector():
    # the timed CPU baseline builds its projector from group orbits: its
    # column count equals the oracle's stage construction for k / p / z sectors
    for L, blocks in ((12, dict(kblock=0, pblock=1, zblock=1)), (12, dict(kblock=3)), (14, dict(zblock=-1)),
                      (12, dict(kblock=6, pblock=-1)), (10, {})):
        ns, t = oracle.bench_qlo_matvec(L, L // 2, False, _dense.xxz_terms(L, J2=0.4), 2, 2, **blocks)
        assert ns == oracle.Sector(L, nup=L // 2, pauli=False, **blocks).Ns
        assert t.shape == (2, 3) and np.all(t >= 0)


@pytest.mark.parametrize("blocks,kw", [
    ([("t", 0), ("p", 0)], dict(kblock=0, pblock=1)),
    ([("t", 0), ("p", 1)], dict(kblock=0, pblock=-1)),
    ([("t", 0), ("z", 0)], dict(kblock=0, zblock=1)),
    ([("t", 3), ("z", 1)], dict(kblock=3, zblock=-1)),
    ([("t", 0), ("p", 1), ("z", 1)], dict(kblock=0, pblock=-1, zblock=-1)),
    ([("t", 1)], dict(kblock=1)),
])
def test_general_maps_reproduce_spin_basis_1d(blocks, kw):
    # SpinBasisGeneral's site maps (basis.jl:282-338) written for translation
    # by two sites, parity and spin inversion give SpinBasis1D's a=2 sectors:
    # same columns, labels and normalizations to the bit (the quantum number q
    # of an order-2 map selects cis(pi q))
    L = 12
    maps = {"t": [(i + 2) % L for i in range(L)], "p": [L - 1 - i for i in range(L)],
            "z": [-(i + 1) for i in range(L)]}
    g = oracle.Sector.general(L, nup=6, pauli=False, blocks=[(maps[n], q) for n, q in blocks])
    a = oracle.Sector(L, nup=6, pauli=False, a=2, **kw)
    assert g.Ns == a.Ns
    assert np.array_equal(g.states(), a.states())
    assert np.array_equal(g.norms()[1].view(np.uint64), a.norms()[1].view(np.uint64))
