"""CPU oracle for the spin-1/2 sector Hamiltonian action (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, as the checker or as the timed CPU baseline.
The product library (``libqsgpu.so``) never loads it.

``qs_oracle.c`` restates QuSpin.jl's CPU algorithms (cited per function there);
``WideSector`` / ``parent_hamiltonian`` below cover chains of up to 256 sites
(basis.jl:361-445, 1231-1300); ``krylov.py`` restates the Krylov exponential
action (operators.jl:1655-1893) and ``rk45.py`` the dynamic evolution
(operators.jl:1964-2071) in numpy.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libqsoracle.so")
UNSET = -1000000
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    c_i64, c_int, c_dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_double
    vp = ctypes.c_void_p
    L.qo_last_error.restype = ctypes.c_char_p
    L.qo_fixed_weight_states.argtypes = [c_int, c_int, vp]
    L.qo_fixed_weight_states.restype = c_i64
    L.qo_sector_build.argtypes = [c_int] * 10
    L.qo_sector_build.restype = vp
    L.qo_sector_build_general.argtypes = [c_int, c_int, c_int, c_int, vp, vp]
    L.qo_sector_build_general.restype = vp
    for name in ("qo_sector_ncols", "qo_sector_nparent", "qo_sector_nnz"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = c_i64
    L.qo_sector_reduced.argtypes = [vp]
    L.qo_sector_reduced.restype = c_int
    for name in ("qo_sector_parent_states", "qo_sector_labels"):
        getattr(L, name).argtypes = [vp, vp]
        getattr(L, name).restype = None
    L.qo_sector_projector.argtypes = [vp, vp, vp, vp]
    L.qo_sector_projector.restype = None
    L.qo_sector_norms.argtypes = [vp, vp, vp]
    L.qo_sector_norms.restype = None
    L.qo_sector_free.argtypes = [vp]
    L.qo_sector_free.restype = None
    L.qo_wsector_build.argtypes = [c_int, c_int, c_int, c_int, vp, vp]
    L.qo_wsector_build.restype = vp
    for name in ("qo_wsector_ncols", "qo_wsector_nparent", "qo_wsector_nnz"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = c_i64
    L.qo_wsector_reduced.argtypes = [vp]
    L.qo_wsector_reduced.restype = c_int
    for name in ("qo_wsector_parent_states", "qo_wsector_labels"):
        getattr(L, name).argtypes = [vp, vp]
        getattr(L, name).restype = None
    L.qo_wsector_projector.argtypes = [vp, vp, vp, vp]
    L.qo_wsector_projector.restype = None
    L.qo_wsector_norms.argtypes = [vp, vp, vp]
    L.qo_wsector_norms.restype = None
    L.qo_wsector_free.argtypes = [vp]
    L.qo_wsector_free.restype = None
    L.qo_op_create.argtypes = [vp, c_int, vp, vp, vp, vp, c_int, vp]
    L.qo_op_create.restype = vp
    L.qo_op_matvec.argtypes = [vp, vp, vp, c_dbl, c_dbl, c_dbl, c_dbl, c_int, c_int]
    L.qo_op_matvec.restype = c_int
    L.qo_op_free.argtypes = [vp]
    L.qo_op_free.restype = None
    L.qo_op_ncouplings.argtypes = [vp]
    L.qo_op_ncouplings.restype = c_int
    L.qo_csr_build.argtypes = [vp, c_int, vp, vp, vp, vp]
    L.qo_csr_build.restype = vp
    L.qo_csr_nnz.argtypes = [vp]
    L.qo_csr_nnz.restype = c_i64
    L.qo_csr_matvec.argtypes = [vp, vp, vp, c_dbl, c_dbl, c_int]
    L.qo_csr_matvec.restype = None
    L.qo_csr_export.argtypes = [vp, vp, vp, vp]
    L.qo_csr_export.restype = None
    L.qo_csr_free.argtypes = [vp]
    L.qo_csr_free.restype = None
    L.qo_bench_parent_apply.argtypes = [c_int, c_int, c_int, c_int, vp, vp, vp, vp, c_i64, c_int]
    L.qo_bench_parent_apply.restype = c_dbl
    L.qo_bench_parent_apply_mt.argtypes = [c_int, c_int, c_int, c_int, vp, vp, vp, vp, c_i64, c_int, c_int]
    L.qo_bench_parent_apply_mt.restype = c_dbl
    L.qo_bench_qlo_matvec.argtypes = [c_int] * 7 + [vp, vp, vp, vp, c_int, c_int, vp]
    L.qo_bench_qlo_matvec.restype = c_i64
    L.qo_prim_hypot.argtypes = [c_dbl, c_dbl]
    L.qo_prim_hypot.restype = c_dbl
    L.qo_prim_cdiv.argtypes = [c_dbl, c_dbl, c_dbl, c_dbl, vp]
    L.qo_prim_cdiv.restype = None
    L.qo_prim_norm2.argtypes = [vp, c_int]
    L.qo_prim_norm2.restype = c_dbl
    _lib = L
    return L


def prim_hypot(x, y):
    """The oracle's |z| (Julia Base.Math.hypot)."""
    return lib().qo_prim_hypot(float(x), float(y))


def prim_cdiv(a, b, c, d):
    """The oracle's complex division (a + ib) / (c + id) (Julia Base, Baudin & Smith 2012)."""
    out = np.zeros(2)
    lib().qo_prim_cdiv(float(a), float(b), float(c), float(d), _ptr(out))
    return complex(out[0], out[1])


def prim_norm2(v):
    """The oracle's LinearAlgebra.generic_norm2 of a complex vector."""
    z = np.ascontiguousarray(np.asarray(v, dtype=np.complex128))
    return lib().qo_prim_norm2(_ptr(z.view(np.float64)), z.size)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def fixed_weight_states(L, w):
    """basis.jl:229-250 (w=None enumerates every L-bit word)."""
    lb = lib()
    ww = -1 if w is None else int(w)
    n = lb.qo_fixed_weight_states(L, ww, None)
    out = np.empty(n, dtype=np.uint64)
    lb.qo_fixed_weight_states(L, ww, _ptr(out))
    return out


def _u(v):
    return UNSET if v is None else int(v)


class Sector:
    """Reference-semantics SpinBasis1D (basis.jl:622-777): explicit projector P."""

    def __init__(self, L, nup=None, pauli=True, a=1, kblock=None, pblock=None, zblock=None,
                 pzblock=None, zAblock=None, zBblock=None):
        lb = lib()
        h = lb.qo_sector_build(L, _u(nup), int(bool(pauli)), a, _u(kblock), _u(pblock), _u(zblock),
                               _u(pzblock), _u(zAblock), _u(zBblock))
        if not h:
            raise ValueError(lb.qo_last_error().decode())
        self._h = h
        self.L, self.nup, self.pauli = L, nup, pauli
        self.Ns = lb.qo_sector_ncols(h)
        self.npar = lb.qo_sector_nparent(h)
        self.reduced = bool(lb.qo_sector_reduced(h))

    @classmethod
    def general(cls, L, nup=None, pauli=True, blocks=()):
        """Reference-semantics SpinBasisGeneral (basis.jl:530-620): `blocks` is a
        sequence of (map, quantum_number) in stage order, each map one entry per
        site in QuSpin's zero-based or Julia's one-based convention."""
        lb = lib()
        maps = np.ascontiguousarray(np.asarray([list(m) for m, _ in blocks] or [[0]], dtype=np.int32))
        qn = np.ascontiguousarray(np.asarray([int(q) for _, q in blocks] or [0], dtype=np.int32))
        h = lb.qo_sector_build_general(L, _u(nup), int(bool(pauli)), len(blocks), _ptr(maps), _ptr(qn))
        if not h:
            raise ValueError(lb.qo_last_error().decode())
        self = cls.__new__(cls)
        self._h = h
        self.L, self.nup, self.pauli = L, nup, pauli
        self.Ns = lb.qo_sector_ncols(h)
        self.npar = lb.qo_sector_nparent(h)
        self.reduced = bool(lb.qo_sector_reduced(h))
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            lib().qo_sector_free(self._h)
            self._h = None

    def parent_states(self):
        out = np.empty(self.npar, dtype=np.uint64)
        lib().qo_sector_parent_states(self._h, _ptr(out))
        return out

    def states(self):
        out = np.empty(self.Ns, dtype=np.uint64)
        lib().qo_sector_labels(self._h, _ptr(out))
        return out

    def norms(self):
        ref = np.empty(self.Ns, dtype=np.float64)
        orb = np.empty(self.Ns, dtype=np.uint32)
        lib().qo_sector_norms(self._h, _ptr(ref), _ptr(orb))
        return orb, ref

    def projector(self):
        """Dense-free CSC pieces (colptr, row, values) of P, 0-based."""
        if not self.reduced:
            return None
        nnz = lib().qo_sector_nnz(self._h)
        colptr = np.empty(self.Ns + 1, dtype=np.int64)
        row = np.empty(nnz, dtype=np.int64)
        val = np.empty(nnz, dtype=np.complex128)
        lib().qo_sector_projector(self._h, _ptr(colptr), _ptr(row), _ptr(val))
        return colptr, row, val

    def projector_dense(self):
        if not self.reduced:
            return np.eye(self.Ns, dtype=np.complex128)
        colptr, row, val = self.projector()
        P = np.zeros((self.npar, self.Ns), dtype=np.complex128)
        for c in range(self.Ns):
            P[row[colptr[c]:colptr[c + 1]], c] = val[colptr[c]:colptr[c + 1]]
        return P


def _words_to_int(w):
    return int(w[0]) | int(w[1]) << 64 | int(w[2]) << 128 | int(w[3]) << 192


class WideSector:
    """Reference-semantics SpinBasisGeneral on up to 256 sites (basis.jl:361-445,
    530-620; the reference's UInt256 states): states are Python ints, the
    parent is the nup sector in ascending order, `blocks` as in Sector.general."""

    def __init__(self, L, nup, pauli=True, blocks=()):
        lb = lib()
        maps = np.ascontiguousarray(np.asarray([list(m) for m, _ in blocks] or [[0]], dtype=np.int32))
        qn = np.ascontiguousarray(np.asarray([int(q) for _, q in blocks] or [0], dtype=np.int32))
        h = lb.qo_wsector_build(L, int(nup), int(bool(pauli)), len(blocks), _ptr(maps), _ptr(qn))
        if not h:
            raise ValueError(lb.qo_last_error().decode())
        self._h = h
        self.L, self.nup, self.pauli = L, nup, pauli
        self.Ns = lb.qo_wsector_ncols(h)
        self.npar = lb.qo_wsector_nparent(h)
        self.reduced = bool(lb.qo_wsector_reduced(h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().qo_wsector_free(self._h)
            self._h = None

    def parent_states(self):
        out = np.empty((self.npar, 4), dtype=np.uint64)
        lib().qo_wsector_parent_states(self._h, _ptr(out))
        return [_words_to_int(r) for r in out]

    def states(self):
        out = np.empty((self.Ns, 4), dtype=np.uint64)
        lib().qo_wsector_labels(self._h, _ptr(out))
        return [_words_to_int(r) for r in out]

    def norms(self):
        ref = np.empty(self.Ns, dtype=np.float64)
        orb = np.empty(self.Ns, dtype=np.uint32)
        lib().qo_wsector_norms(self._h, _ptr(ref), _ptr(orb))
        return orb, ref

    def projector_sparse(self):
        """P as a scipy CSC matrix (npar x Ns); the identity without blocks."""
        import scipy.sparse as sp
        if not self.reduced:
            return sp.identity(self.npar, dtype=np.complex128, format="csc")
        nnz = lib().qo_wsector_nnz(self._h)
        colptr = np.empty(self.Ns + 1, dtype=np.int64)
        row = np.empty(nnz, dtype=np.int64)
        val = np.empty(nnz, dtype=np.complex128)
        lib().qo_wsector_projector(self._h, _ptr(colptr), _ptr(row), _ptr(val))
        return sp.csc_matrix((val, row, colptr), shape=(self.npar, self.Ns))


def _local(op, state, site, pauli):
    """basis.jl:1231-1259 (_apply_wide_spin_local) / operators.jl:343-379:
    (new state, factor, alive) of one local operator on 1-based `site`."""
    mask = 1 << (site - 1)
    occ = bool(state & mask)
    half = 1.0 if pauli else 0.5
    if op == "I":
        return state, 1.0, True
    if op == "z":
        return state, (half if occ else -half), True
    if op == "+":
        if occ:
            return state, 0.0, False
        return state | mask, (2.0 if pauli else 1.0), True
    if op == "-":
        if not occ:
            return state, 0.0, False
        return state & ~mask, (2.0 if pauli else 1.0), True
    if op == "x":
        return state ^ mask, half, True
    if op == "y":
        return state ^ mask, (-1j * half if occ else 1j * half), True
    raise ValueError("unsupported spin operator %r" % op)


def parent_hamiltonian(states, terms, pauli=True):
    """basis.jl:1261-1300 (_wide_spin_operator_triplets): the operator list on
    the parent states (ascending Python ints) as a scipy CSC matrix; operator
    strings act right to left, couplings (coef, site1, site2, ...) 1-based."""
    import scipy.sparse as sp
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for op, couplings in terms:
        for c in couplings:
            sites = c[1:]
            for col, init in enumerate(states):
                st, amp, alive = init, complex(c[0]), True
                for k in range(len(op) - 1, -1, -1):
                    st, f, alive = _local(op[k], st, int(sites[k]), pauli)
                    if not alive:
                        break
                    amp *= f
                if not alive:
                    continue
                r = index.get(st)
                if r is None:
                    continue
                rows.append(r)
                cols.append(col)
                vals.append(amp)
    n = len(states)
    return sp.csc_matrix((np.asarray(vals, dtype=np.complex128), (rows, cols)), shape=(n, n))


def wide_sector_matrix(sector, terms):
    """P' H P of a WideSector (the projected triplet matrix of basis.jl:1302-1326)."""
    P = sector.projector_sparse()
    H = parent_hamiltonian(sector.parent_states(), terms, sector.pauli)
    return (P.conj().T @ H @ P).toarray()


def _flatten_terms(terms):
    ops, ncoup, coef, sites = [], [], [], []
    for op, couplings in terms:
        ops.append(op.encode())
        ncoup.append(len(couplings))
        for c in couplings:
            z = complex(c[0])
            coef += [z.real, z.imag]
            sites += [int(s) for s in c[1:]]
    arr_ops = (ctypes.c_char_p * max(1, len(ops)))(*ops)
    return (arr_ops, np.asarray(ncoup, dtype=np.int32), np.asarray(coef, dtype=np.float64),
            np.asarray(sites if sites else [0], dtype=np.int32))


class Operator:
    """Reference-semantics QuantumLinearOperator (operators.jl:2492-2552)."""

    def __init__(self, sector, terms, check_symm=True, diagonal=None):
        lb = lib()
        self.sector = sector
        self._keep = _flatten_terms(terms)
        ops, ncoup, coef, sites = self._keep
        d = None
        if diagonal is not None:
            self._diag = np.ascontiguousarray(np.asarray(diagonal, dtype=np.complex128))
            d = _ptr(self._diag)
        h = lb.qo_op_create(sector._h, len(terms), ops, _ptr(ncoup), _ptr(coef), _ptr(sites),
                            int(bool(check_symm)), d)
        if not h:
            raise ValueError(lb.qo_last_error().decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().qo_op_free(self._h)
            self._h = None

    def matvec(self, x, y=None, alpha=1.0, beta=0.0, transposed=False, conjugated=False):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
        out = np.zeros(self.sector.Ns, dtype=np.complex128) if y is None else \
            np.ascontiguousarray(np.asarray(y, dtype=np.complex128)).copy()
        a, b = complex(alpha), complex(beta)
        lib().qo_op_matvec(self._h, _ptr(x), _ptr(out), a.real, a.imag, b.real, b.imag,
                           int(transposed), int(conjugated))
        return out

    def dense(self):
        n = self.sector.Ns
        M = np.zeros((n, n), dtype=np.complex128)
        e = np.zeros(n, dtype=np.complex128)
        for j in range(n):
            e[:] = 0
            e[j] = 1
            M[:, j] = self.matvec(e)
        return M


class CSR:
    """Assembled real CSR Hamiltonian on a basis without symmetry
    (operators.jl:381-576, sparse_formats.jl:29-50,110-146)."""

    def __init__(self, sector, terms):
        lb = lib()
        self._keep = _flatten_terms(terms)
        ops, ncoup, coef, sites = self._keep
        h = lb.qo_csr_build(sector._h, len(terms), ops, _ptr(ncoup), _ptr(coef), _ptr(sites))
        if not h:
            raise ValueError(lb.qo_last_error().decode())
        self._h = h
        self.n = sector.Ns
        self.nnz = lb.qo_csr_nnz(h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().qo_csr_free(self._h)
            self._h = None

    def matvec(self, x, y=None, alpha=1.0, beta=0.0, nthreads=1):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        out = np.zeros(self.n) if y is None else y
        lib().qo_csr_matvec(self._h, _ptr(x), _ptr(out), float(alpha), float(beta), int(nthreads))
        return out

    def export(self):
        rowptr = np.empty(self.n + 1, dtype=np.int64)
        col = np.empty(self.nnz, dtype=np.int64)
        val = np.empty(self.nnz, dtype=np.float64)
        lib().qo_csr_export(self._h, _ptr(rowptr), _ptr(col), _ptr(val))
        return rowptr, col, val


def bench_parent_apply(L, nup, pauli, terms, n_sample, repeats=1):
    """Seconds the reference's parent-space apply loop takes over the first
    n_sample parent states (single thread); see qo_bench_parent_apply."""
    ops, ncoup, coef, sites = _flatten_terms(terms)
    return lib().qo_bench_parent_apply(L, nup, int(bool(pauli)), len(terms), ops, _ptr(ncoup), _ptr(coef),
                                       _ptr(sites), int(n_sample), int(repeats))


def bench_parent_apply_mt(L, nup, pauli, terms, n_sample, threads, repeats=1):
    """The same loop on `threads` OpenMP threads (columns split, atomic scatter)."""
    ops, ncoup, coef, sites = _flatten_terms(terms)
    return lib().qo_bench_parent_apply_mt(L, nup, int(bool(pauli)), len(terms), ops, _ptr(ncoup), _ptr(coef),
                                          _ptr(sites), int(n_sample), int(repeats), int(threads))


def bench_qlo_matvec(L, nup, pauli, terms, threads, repeats, kblock=None, pblock=None, zblock=None):
    """Seconds of `repeats` full products of the reference's QLO (operators.jl:2591-2623: lift, parent
    apply with a hash lookup per hit, project) on `threads` threads: (Ns, array of shape (repeats, 3)
    with the lift / apply / project seconds)."""
    keep = _flatten_terms(terms)
    ops, ncoup, coef, sites = keep
    t = np.zeros((repeats, 3))
    ns = lib().qo_bench_qlo_matvec(L, _u(nup), int(bool(pauli)), _u(kblock), _u(pblock), _u(zblock), len(terms),
                                   ops, _ptr(ncoup), _ptr(coef), _ptr(sites), int(threads), int(repeats), _ptr(t))
    if ns < 0:
        raise MemoryError("qo_bench_qlo_matvec failed")
    return int(ns), t
