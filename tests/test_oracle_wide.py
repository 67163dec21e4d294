"""Pins the oracle's wide sector (states of up to 256 sites, basis.jl:361-445)
and its Python parent Hamiltonian before they are trusted as the checker.

Below 64 sites the wide builder must agree bit for bit with the oracle's
SpinBasisGeneral sector, which tests/test_oracle.py pins against SpinBasis1D;
the Python parent Hamiltonian must agree with the oracle's operator; above 64
sites, where no other construction exists, closed forms fix the result: the
one-magnon dispersion of the XXZ ring and the free-fermion spectrum of the
two-particle XX ring, which every momentum sector together must reproduce.
"""
import math

import numpy as np
import pytest

import oracle
import _dense


def _maps(L):
    return {"t": [(i + 1) % L for i in range(L)], "t2": [(i + 2) % L for i in range(L)],
            "p": [L - 1 - i for i in range(L)], "z": [-(i + 1) for i in range(L)]}


SECTORS = [
    (12, 6, [("t", 0)]), (12, 6, [("t", 5)]), (12, 6, [("p", 1)]), (12, 6, [("z", 0)]),
    (12, 6, [("t2", 0), ("p", 0)]), (12, 6, [("t", 0), ("z", 1)]), (12, 6, [("t2", 3), ("z", 0)]),
    (11, 4, [("t", 2)]), (13, 3, [("p", 0)]), (10, 5, [("t", 0), ("p", 0), ("z", 0)]),
]


@pytest.mark.parametrize("L,nup,blocks", SECTORS)
def test_wide_sector_equals_general_sector(L, nup, blocks):
    m = _maps(L)
    bl = [(m[n], q) for n, q in blocks]
    w = oracle.WideSector(L, nup, pauli=False, blocks=bl)
    g = oracle.Sector.general(L, nup=nup, pauli=False, blocks=bl)
    assert w.Ns == g.Ns and w.npar == g.npar
    assert w.parent_states() == [int(s) for s in g.parent_states()]
    assert w.states() == [int(s) for s in g.states()]
    ow, rw = w.norms()
    og, rg = g.norms()
    assert np.array_equal(ow, og)
    assert np.array_equal(rw.view(np.uint64), rg.view(np.uint64))
    Pw = w.projector_sparse().toarray()
    assert np.array_equal(Pw, g.projector_dense())


def test_wide_sector_without_blocks_is_the_particle_sector():
    w = oracle.WideSector(9, 4, pauli=True)
    assert not w.reduced and w.Ns == math.comb(9, 4)
    assert w.states() == [s for s in range(2 ** 9) if bin(s).count("1") == 4]


@pytest.mark.parametrize("L,nup,blocks", SECTORS[:7])
def test_parent_hamiltonian_against_oracle_operator(L, nup, blocks):
    m = _maps(L)
    bl = [(m[n], q) for n, q in blocks]
    terms = _dense.xxz_terms(L, J=1.0, delta=0.7, J2=0.3)
    w = oracle.WideSector(L, nup, pauli=False, blocks=bl)
    got = oracle.wide_sector_matrix(w, terms)
    want = oracle.Operator(oracle.Sector.general(L, nup=nup, pauli=False, blocks=bl), terms).dense()
    assert np.abs(got - want).max() <= 1e-13


def test_parent_hamiltonian_y_and_pauli_conventions():
    # 'y' and the Pauli scale against the dense Kronecker construction
    L = 6
    terms = [("xy", [(0.5, i, i % L + 1) for i in range(1, L + 1)]), ("z", [(0.3, 2)]), ("+-", [(0.25, 1, 4)])]
    for pauli in (True, False):
        w = oracle.WideSector(L, 3, pauli=pauli)
        got = oracle.parent_hamiltonian(w.parent_states(), terms, pauli).toarray()
        full = _dense.full_hamiltonian(L, terms, pauli)
        idx = w.parent_states()
        assert np.abs(got - full[np.ix_(idx, idx)]).max() <= 1e-14


def test_one_magnon_dispersion_on_seventy_sites():
    # one flipped spin on an XXZ ring of 70 sites: every momentum sector holds
    # one state of energy J cos(2 pi k / L) + Delta (L - 4) / 4 (spin-1/2 operators)
    L, J, delta = 70, 1.0, 0.6
    terms = _dense.xxz_terms(L, J=J, delta=delta)
    t = _maps(L)["t"]
    for k in (0, 1, 17, 35, 69):
        w = oracle.WideSector(L, 1, pauli=False, blocks=[(t, k)])
        # the label is the first state of largest |P| entry, which rounding picks
        assert w.Ns == 1 and bin(w.states()[0]).count("1") == 1
        e = oracle.wide_sector_matrix(w, terms)[0, 0]
        assert abs(e - (J * math.cos(2 * math.pi * k / L) + delta * (L - 4) / 4)) <= 1e-12


def test_two_particle_xx_ring_spectrum_on_seventy_sites():
    # two particles on an XX ring of 70 sites are free fermions with
    # antiperiodic momenta (Jordan-Wigner, even particle number): the spectra
    # of all 70 momentum sectors together are the pair energies
    # J (cos k1 + cos k2), k = 2 pi (m + 1/2) / L, m1 < m2
    L, J = 70, 1.0
    terms = _dense.xxz_terms(L, J=J, delta=0.0)
    t = _maps(L)["t"]
    first = oracle.WideSector(L, 2, pauli=False, blocks=[(t, 0)])
    states = first.parent_states()
    assert max(states).bit_length() == 70      # states span both 64-bit words
    H = oracle.parent_hamiltonian(states, terms, False)
    got = []
    dims = []
    for k in range(L):
        w = first if k == 0 else oracle.WideSector(L, 2, pauli=False, blocks=[(t, k)])
        P = w.projector_sparse()
        M = (P.conj().T @ H @ P).toarray()
        assert np.abs(M - M.conj().T).max() <= 1e-13
        dims.append(w.Ns)
        got.extend(np.linalg.eigvalsh(M))
    assert sum(dims) == math.comb(L, 2)
    assert dims[0] == 35 and dims[1] == 34       # the distance-35 orbit needs an even k
    ks = [2 * math.pi * (m + 0.5) / L for m in range(L)]
    want = sorted(J * (math.cos(ks[a]) + math.cos(ks[b])) for a in range(L) for b in range(a + 1, L))
    assert np.abs(np.sort(got) - np.asarray(want)).max() <= 1e-11


def test_wide_sector_rejects_bad_input():
    with pytest.raises(ValueError):
        oracle.WideSector(300, 2)
    with pytest.raises(ValueError):
        oracle.WideSector(70, 2, blocks=[([0] * 70, 0)])
