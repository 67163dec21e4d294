"""Pins of the oracle's floating-point primitives, labels and Krylov restatement.

The oracle repeats the reference's floating-point sequence to reproduce its
representative labels and normalizations (symmetry_basis.jl:57-232,
discrete_basis.jl:2386-2422).  These checks tie each piece to something the
oracle does not compute itself:
  * its |z| (Julia's hypot), complex division (Baudin & Smith) and
    generic_norm2 against exact rational arithmetic;
  * the labels and normalizations of composite sectors against properties
    the reference's definitions fix (the label lies in its column's orbit and
    is the first row of largest |P|, normalization = 1/abs2(P[label, c]) to
    the bit and within 1e-14 relative of the orbit size);
  * the prediction, from the reference's stage construction, that a k=0
    parity sector whose orbit lengths all divide a power of two labels every
    column with its orbit minimum;
  * the numpy Krylov restatement (oracle/krylov.py) against dense
    exponentials and the reference's own known answers.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from oracle import krylov
import _dense

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))
EPS = np.finfo(np.float64).eps


# ------------------------------------------------------------ exact helpers --
def _next_up(x):
    return float(np.nextafter(x, np.inf))


def _exact_sqrt_rounded(S):
    """Nearest double to sqrt(S) for a nonnegative Fraction S (ties to even)."""
    if S == 0:
        return 0.0
    h = math.sqrt(float(S))
    # move to the largest double h1 with h1^2 <= S
    while Fraction(h) ** 2 > S:
        h = float(np.nextafter(h, 0.0))
    while Fraction(_next_up(h)) ** 2 <= S:
        h = _next_up(h)
    h2 = _next_up(h)
    mid = (Fraction(h) + Fraction(h2)) / 2
    if S < mid * mid:
        return h
    if S > mid * mid:
        return h2
    return h if (np.frombuffer(np.float64(h).tobytes(), np.uint64)[0] & 1) == 0 else h2


def _ulp(x):
    return float(np.spacing(abs(x)))


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        e1, e2 = rng.integers(-40, 40, 2)
        out.append((float(rng.standard_normal() * 2.0 ** e1), float(rng.standard_normal() * 2.0 ** e2)))
    # the magnitudes the projector stages produce: 1/sqrt(R), phases of roots of unity
    for R in range(1, 65):
        a = 1 / math.sqrt(R)
        for k in range(min(R, 6)):
            out.append((a * math.cos(2 * math.pi * k / R), a * math.sin(2 * math.pi * k / R)))
    return out


# --------------------------------------------------------------- primitives --
def test_hypot_is_correctly_rounded():
    # Base.Math.hypot on FMA hardware: the oracle's |z| must be the nearest
    # double to sqrt(x^2 + y^2) (a dropped correction step is off by an ulp)
    bad = []
    for x, y in _samples(3000, 11):
        got = oracle.prim_hypot(x, y)
        want = _exact_sqrt_rounded(Fraction(x) ** 2 + Fraction(y) ** 2)
        if got != want:
            bad.append((x, y, got, want))
    assert not bad, bad[:5]


def test_complex_division_against_exact_quotient():
    # Baudin & Smith's division is accurate to a few ulps normwise; swapped
    # operands, a lost sign or a wrong branch are off by O(1)
    for (a, b), (c, d) in zip(_samples(1500, 12), _samples(1500, 13)[::-1]):
        if c == 0 and d == 0:
            continue
        got = oracle.prim_cdiv(a, b, c, d)
        den = Fraction(c) ** 2 + Fraction(d) ** 2
        qr = (Fraction(a) * Fraction(c) + Fraction(b) * Fraction(d)) / den
        qi = (Fraction(b) * Fraction(c) - Fraction(a) * Fraction(d)) / den
        mag = max(abs(float(qr)), abs(float(qi)))
        assert abs(Fraction(got.real) - qr) <= 4 * EPS * Fraction(mag), (a, b, c, d, got)
        assert abs(Fraction(got.imag) - qi) <= 4 * EPS * Fraction(mag), (a, b, c, d, got)


def test_generic_norm2_against_exact_norm():
    rng = np.random.default_rng(14)
    for n in (1, 2, 3, 7, 12, 28, 31):
        for _ in range(40):
            v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            got = oracle.prim_norm2(v)
            S = sum(Fraction(z.real) ** 2 + Fraction(z.imag) ** 2 for z in v)
            want = _exact_sqrt_rounded(S)
            assert abs(got - want) <= (n + 1) * _ulp(want), (n, got, want)


# ---------------------------------------------------- labels and norms ------
COMPOSITE = [
    (8, 4, dict(kblock=0, pblock=1)), (8, 4, dict(kblock=4, pblock=-1, zblock=1)),
    (10, 5, dict(kblock=0, pblock=1, zblock=-1)), (10, 5, dict(kblock=5, pblock=1)),
    (12, 6, dict(kblock=0, pblock=1, zblock=1)), (12, 6, dict(kblock=6, pblock=-1, zblock=-1)),
    (12, 4, dict(kblock=0, pblock=1)), (12, 6, dict(kblock=2, zblock=1)), (12, 6, dict(kblock=3)),
    (14, 7, dict(kblock=0, pblock=1, zblock=1)), (14, 7, dict(kblock=7, pblock=1, zblock=-1)),
    (16, 8, dict(kblock=0, pblock=1, zblock=1)), (16, 8, dict(kblock=8, zblock=1)),
    (12, 6, dict(pblock=1, zblock=1)), (12, 6, dict(a=2, kblock=1)), (12, 6, dict(kblock=0, pzblock=1)),
]


def _columns(sec):
    colptr, row, val = sec.projector()
    parent = sec.parent_states()
    for c in range(sec.Ns):
        sl_ = slice(colptr[c], colptr[c + 1])
        yield c, parent[row[sl_]], val[sl_]


@pytest.mark.parametrize("L,nup,blocks", COMPOSITE)
def test_labels_and_normalizations_follow_their_definitions(L, nup, blocks):
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    labels = sec.states()
    orbit, ref = sec.norms()
    group = [g for g, _ in _dense.group_elements(L, **blocks)]
    for c, states, vals in _columns(sec):
        lab = int(labels[c])
        # the column's rows are one orbit of the group, listed in ascending
        # parent order, and the label is one of them
        assert lab in states.tolist()
        orb = {int(g(int(states[0]))) for g in group}
        assert set(states.tolist()) <= orb
        assert list(states) == sorted(states)
        # first row of largest |P| (strict > scan, symmetry_basis.jl:224-228)
        mags = [oracle.prim_hypot(z.real, z.imag) for z in vals]
        best = 0
        for i in range(1, len(mags)):
            if mags[i] > mags[best]:
                best = i
        assert lab == int(states[best])
        # normalization = inv(abs2(amp)) at the label's row, to the bit
        amp = vals[best]
        assert ref[c] == 1.0 / (amp.real * amp.real + amp.imag * amp.imag)
        # and the orbit size D up to rounding of the stage arithmetic (the
        # stages multiply and divide by norms and overlaps that are 1 only up
        # to rounding: a few tens of ulps after three stages)
        assert orbit[c] == len(states)
        assert abs(ref[c] - len(states)) <= 1e-14 * len(states), (c, ref[c], len(states))


@pytest.mark.parametrize("L", [8, 16])
@pytest.mark.parametrize("p", [1, -1])
def test_k0_parity_labels_are_orbit_minima_when_orbits_divide_a_power_of_two(L, p):
    # For k=0 every stage-1 entry of an orbit of R states is 1/sqrt(R); the
    # parity stage combines column a with partner b through their overlap
    # phi = sum fl(x x), which is exactly 1 or just below 1 for R in {1,2,4,8,16}.
    # Entries of a then stay the largest, so the label is the union minimum.
    sec = oracle.Sector(L, nup=L // 2, pauli=False, kblock=0, pblock=p)
    _, _, mins = _dense.sector_projector(L, L // 2, kblock=0, pblock=p)
    assert sec.states().tolist() == mins.tolist()


def test_k0_parity_labels_leave_the_orbit_minimum_for_other_orbit_lengths():
    # ... while orbit lengths 3, 6 and 12 give phi slightly above 1 and the
    # reference labels some columns with the partner orbit's minimum
    sec = oracle.Sector(12, nup=6, pauli=False, kblock=0, pblock=1)
    _, _, mins = _dense.sector_projector(12, 6, kblock=0, pblock=1)
    labels = sec.states()
    assert sorted(set(labels.tolist()) | set(mins.tolist())) != sorted(mins.tolist())
    assert len(labels) == len(mins)


# ---------------------------------------------------- Krylov restatement ----
def test_krylov_restatement_rotation_known_answer():
    # ExpmMultiplyParallel on the 2x2 generator (runtests.jl:198-213)
    g = GOLD["expm_parallel_rotation"]
    G = np.array(g["generator"], dtype=np.complex128)
    out = krylov.expmv_vector(G, np.array(g["v"], dtype=np.complex128), g["a"], hermitian=False)
    assert np.allclose(out, g["expected"], atol=1e-14)


def test_krylov_restatement_against_dense_exponentials():
    g = GOLD["matrix_free_krylov_L8"]
    L = g["L"]
    terms = [("zz", [(g["zz"], i, i % L + 1) for i in range(1, L + 1)]),
             ("+-", [(g["pm"], i, i % L + 1) for i in range(1, L + 1)]),
             ("-+", [(g["pm"], i, i % L + 1) for i in range(1, L + 1)])]
    sec = oracle.Sector(L, nup=g["nup"], pauli=g["pauli"])
    H = oracle.Operator(sec, terms).dense()
    rng = np.random.default_rng(3)
    v = rng.standard_normal(sec.Ns) + 1j * rng.standard_normal(sec.Ns)
    v /= np.linalg.norm(v)
    scales = [-1j * t for t in g["times"]]
    got = krylov.expmv_times(H, v, scales, hermitian=True)
    for j, s in enumerate(scales):
        assert np.allclose(got[:, j], sl.expm(s * H) @ v, atol=g["atol_evolve"])
    s = 1j * g["expm_scale_im"]
    assert np.allclose(krylov.expmv_vector(H, v, s, hermitian=True), sl.expm(s * H) @ v, atol=g["atol_evolve"])
    # the Arnoldi path (two-pass Gram-Schmidt) on a non-Hermitian matrix
    A = H + 0.3 * np.triu(H, 1)
    want = sl.expm(-0.6j * A) @ v
    assert np.allclose(krylov.expmv_vector(A, v, -0.6j, hermitian=False), want, atol=1e-11)
    # long times take the halving recursion
    want = sl.expm(-6j * H) @ v
    assert np.allclose(krylov.expmv_vector(H, v, -6j, hermitian=True, krylov_dim=12), want, atol=1e-9)


def test_krylov_restatement_reproduces_the_lanczos_orthogonality_loss():
    # On a sector smaller than krylov_dim the reference's Hermitian path runs
    # Lanczos without reorthogonalisation to the full dimension
    # (operators.jl:1686-1700) and its result drifts from exp(-iHt)v: on the
    # L=10 k=8 evolve request of the serving test stream (25 states, t = 9.14)
    # the drift is about 1.1e-8.  The serving test bounds the GPU by it.
    from qsgpu import serve
    req = [r for r in serve.poisson_workload(14, rate=200.0, seed=7, Ls=(10, 12))
           if r["kind"] == "evolve" and r["blocks"] == {"kblock": 8}][0]
    sec = oracle.Sector(req["L"], nup=req["nup"], pauli=False, **req["blocks"])
    H = oracle.Operator(sec, serve.chain_terms(req["L"], req["J2"], req["delta"]), check_symm=False).dense()
    rng = np.random.default_rng(req["seed"])
    v = rng.standard_normal(H.shape[0]) + 0j
    v /= np.linalg.norm(v)
    got = krylov.expmv_times(H, v, [-1j * req["t"]], hermitian=True)[:, 0]
    err = np.abs(got - sl.expm(-1j * req["t"] * H) @ v).max()
    assert 5e-9 < err < 2e-8, err


# ------------------------------------------------------- RK45 restatement ----
def _driven_chain(L=8):
    from oracle import rk45  # noqa: F401
    sec = oracle.Sector(L, nup=L // 2, pauli=False)
    H0 = oracle.Operator(sec, _dense.xxz_terms(L, J=1.0, delta=0.6)).dense()
    stag = [("z", [((-1.0) ** i * 0.5, i) for i in range(1, L + 1)])]
    H1 = oracle.Operator(sec, stag).dense()
    return H0, H1


def test_rk45_restatement_against_an_independent_integrator():
    # operators.jl:1964-2071 restated; against scipy's DOP853 at tight
    # tolerances for a driven staggered field f(t) = cos(2t)
    from scipy.integrate import solve_ivp
    from oracle import rk45
    H0, H1 = _driven_chain()
    rng = np.random.default_rng(21)
    v = rng.standard_normal(H0.shape[0]) + 1j * rng.standard_normal(H0.shape[0])
    v /= np.linalg.norm(v)
    f = lambda t, w: np.cos(w * t)
    times = [0.5, 1.3, 2.0]
    got = rk45.evolve(H0, [(H1, f, (2.0,))], v, 0.0, times, max_step=0.05)
    sol = solve_ivp(lambda t, y: -1j * (H0 @ y + f(t, 2.0) * (H1 @ y)), (0.0, 2.0), v, method="DOP853",
                    t_eval=times, rtol=1e-13, atol=1e-13)
    assert np.abs(got - sol.y).max() < 2e-8
    # a constant drive is the static exponential of H0 + c H1
    got = rk45.evolve(H0, [(H1, lambda t: 0.7, ())], v, 0.0, [1.0], max_step=0.1)[:, 0]
    assert np.abs(got - sl.expm(-1j * (H0 + 0.7 * H1)) @ v).max() < 1e-8
    # imaginary time renormalises each state
    got = rk45.evolve(H0, [(H1, f, (2.0,))], v, 0.0, [0.4, 0.8], imag_time=True)
    assert np.allclose(np.linalg.norm(got, axis=0), 1.0, atol=1e-14)
