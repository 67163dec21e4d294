"""Independent dense constructions used to pin the oracle and the GPU path.

Nothing here follows the reference's algorithms: operators are Kronecker
products of 2x2 site matrices, and symmetry sectors are built from the group
projector  P_chi |s> = sum_g conj(chi(g)) g|s>  on orbit minima, normalised,
in ascending order of the orbit minimum (the column order QuSpin.jl's stage
construction produces).  Only small L (<= 12) is meant.
"""
import itertools

import numpy as np


def site_matrix(op, pauli):
    """2x2 matrix of one local operator in the basis (bit 0 = down, bit 1 = up).

    Conventions of QuSpin.jl's _apply_local (operators.jl:343-379)."""
    sz = 1.0 if pauli else 0.5
    spm = 2.0 if pauli else 1.0
    M = np.zeros((2, 2), dtype=np.complex128)
    if op == "I":
        M[0, 0] = M[1, 1] = 1
    elif op == "z":
        M[0, 0], M[1, 1] = -sz, sz
    elif op == "+":
        M[1, 0] = spm
    elif op == "-":
        M[0, 1] = spm
    elif op == "x":
        M[0, 1] = M[1, 0] = sz
    elif op == "y":
        M[1, 0] = 1j * sz
        M[0, 1] = -1j * sz
    else:
        raise ValueError(op)
    return M


def embed(M, site, L):
    """Operator acting with M on (1-based) site, identity elsewhere; index = state bits."""
    out = np.ones((1, 1), dtype=np.complex128)
    for bit in range(L - 1, -1, -1):
        out = np.kron(out, M if bit == site - 1 else np.eye(2))
    return out


def full_hamiltonian(L, terms, pauli):
    """Dense 2^L x 2^L matrix of sum_coupling coef * A_1 A_2 ... A_n."""
    H = np.zeros((2 ** L, 2 ** L), dtype=np.complex128)
    for op, couplings in terms:
        for c in couplings:
            O = np.eye(2 ** L, dtype=np.complex128) * complex(c[0])
            for ch, site in zip(op, c[1:]):
                O = O @ embed(site_matrix(ch, pauli), site, L)
            H += O
    return H


def _rotl(u, k, L):
    k %= L
    return ((u << k) | (u >> (L - k))) & ((1 << L) - 1) if k else u


def _rev(u, L):
    return int(format(u, "0%db" % L)[::-1], 2)


def group_elements(L, a=1, kblock=None, pblock=None, zblock=None, pzblock=None):
    """List of (map, character) for the abelian-character group of a spin chain."""
    full = (1 << L) - 1
    nT = L // a if kblock is not None else 1
    lam = np.exp(2j * np.pi * (kblock or 0) / nT) if kblock is not None else 1.0
    points = [(lambda u: u, 1.0)]
    if pblock is not None:
        points = points + [(lambda u, f=f: _rev(f(u), L), ch * pblock) for f, ch in points]
    if zblock is not None:
        points = points + [(lambda u, f=f: f(u) ^ full, ch * zblock) for f, ch in points]
    if pzblock is not None:
        points = points + [(lambda u, f=f: _rev(f(u), L) ^ full, ch * pzblock) for f, ch in points]
    out = []
    for t in range(nT):
        for f, ch in points:
            out.append((lambda u, f=f, t=t: _rotl(f(u), t * a, L), ch * lam ** t))
    return out


def sector_projector(L, nup, **blocks):
    """Columns P_chi|s0>/norm over orbit minima s0 of the nup sector (ascending)."""
    states = [s for s in range(2 ** L) if nup is None or bin(s).count("1") == nup]
    index = {s: i for i, s in enumerate(states)}
    G = group_elements(L, **blocks)
    cols, mins = [], []
    seen = set()
    for s in states:
        orbit = {g(s) for g, _ in G}
        m = min(orbit)
        if m in seen:
            continue
        seen.add(m)
        v = np.zeros(len(states), dtype=np.complex128)
        for g, ch in G:
            v[index[g(m)]] += np.conj(ch)
        nrm = np.linalg.norm(v)
        if nrm < 1e-9:
            continue
        cols.append(v / nrm)
        mins.append(m)
    P = np.array(cols).T if cols else np.zeros((len(states), 0))
    return np.array(states, dtype=np.uint64), P, np.array(mins, dtype=np.uint64)


def sector_hamiltonian(L, nup, terms, pauli, **blocks):
    states, P, mins = sector_projector(L, nup, **blocks)
    Hf = full_hamiltonian(L, terms, pauli)
    Hs = Hf[np.ix_(states.astype(np.int64), states.astype(np.int64))]
    return P.conj().T @ Hs @ P, P, mins


def xxz_terms(L, J=1.0, delta=1.0, J2=0.0, periodic=True):
    """J/2 (S+S- + h.c.) + J*delta SzSz on nearest (and J2 next-nearest) bonds."""
    bonds1 = [(i, i % L + 1) for i in range(1, L + 1)] if periodic else [(i, i + 1) for i in range(1, L)]
    terms = [("+-", [(J / 2, i, j) for i, j in bonds1]), ("-+", [(J / 2, i, j) for i, j in bonds1]),
             ("zz", [(J * delta, i, j) for i, j in bonds1])]
    if J2:
        bonds2 = [(i, (i + 1) % L + 1) for i in range(1, L + 1)]
        terms += [("+-", [(J2 / 2, i, j) for i, j in bonds2]), ("-+", [(J2 / 2, i, j) for i, j in bonds2]),
                  ("zz", [(J2 * delta, i, j) for i, j in bonds2])]
    return terms


def burnside_count(L, nup, **blocks):
    """Sector dimension from the character formula dim = (1/|G|) sum_g chi(g)^* |Fix(g)|."""
    G = group_elements(L, **blocks)
    states = [s for s in range(2 ** L) if nup is None or bin(s).count("1") == nup]
    tot = 0
    for g, ch in G:
        fix = sum(1 for s in states if g(s) == s)
        tot += np.conj(ch) * fix
    return int(round((tot / len(G)).real))


def all_sectors_L(L):
    """Every (kblock, pblock, zblock) combination the reference accepts at nup=L/2."""
    out = []
    for k in [None] + list(range(L)):
        ps = [None, 1, -1] if k in (None, 0, L // 2) else [None]
        for p, z in itertools.product(ps, [None, 1, -1]):
            out.append(dict(kblock=k, pblock=p, zblock=z))
    return out
