"""SpinBasis1D and SpinBasisGeneral on the GPU (basis.jl:622-777, 530-620)."""
import ctypes

import numpy as np

from . import _lib
from .errors import ArgumentError, check


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _blk(v):
    return _lib.QS_UNSET if v is None else int(v)


class SpinBasis1D:
    """Spin-1/2 chain basis in a particle-number and symmetry sector.

    Arguments follow SpinBasis1D(L; nup, Nup, m, S, pauli, a, kblock, pblock,
    zblock, pzblock, zAblock, zBblock) in basis.jl:622-639.  The sector is
    enumerated on the GPU; ``states`` are the reference's representative
    labels in column order and ``reps`` the orbit-minimum keys the kernels use.
    """

    def __init__(self, L, nup=None, Nup=None, m=None, S="1/2", pauli=True, a=1, kblock=None, pblock=None,
                 zblock=None, pzblock=None, zAblock=None, zBblock=None):
        if str(S) not in ("1/2", "0.5"):
            raise ArgumentError("only spin-1/2 chains are on the GPU path")
        if nup is not None and Nup is not None:
            raise ArgumentError("specify only one of nup and Nup")
        sel = Nup if nup is None else nup
        if m is not None:
            if sel is not None:
                raise ArgumentError("m cannot be combined with nup or Nup")
            sel = int(round((float(m) + 0.5) * L))
        self.L = int(L)
        self.nup = None if sel is None else int(sel)
        self.pauli = bool(pauli)
        self.a = int(a)
        self._blocks = dict(kblock=kblock, pblock=pblock, zblock=zblock, pzblock=pzblock, zAblock=zAblock,
                            zBblock=zBblock)
        h = ctypes.c_void_p()
        check(_lib.lib().qs_basis_create(self.L, _blk(self.nup), int(self.pauli), self.a, _blk(kblock), _blk(pblock),
                                          _blk(zblock), _blk(pzblock), _blk(zAblock), _blk(zBblock), 0,
                                          ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        check(_lib.lib().qs_basis_size(self._h, ctypes.byref(n)))
        self.Ns = int(n.value)
        c = ctypes.c_int()
        check(_lib.lib().qs_basis_is_complex(self._h, ctypes.byref(c)))
        self.is_complex = bool(c.value)
        self._states = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().qs_basis_destroy(h)
            except (TypeError, AttributeError):
                pass   # interpreter shutdown: module globals are already gone
            self._h = None

    # basis.jl:779-812 -------------------------------------------------------
    def __len__(self):
        return self.Ns

    @property
    def N(self):
        return self.L

    @property
    def blocks(self):
        out = {"nup": self.nup, "pauli": self.pauli}
        for k, v in self._blocks.items():
            if v is not None:
                out[k] = int(v)
        if self._blocks["kblock"] is not None:
            out["a"] = self.a
            out["kblock"] = int(self._blocks["kblock"]) % (self.L // self.a)
        return out

    def _labels(self):
        if self._states is None:
            out = np.empty(self.Ns, dtype=np.uint64)
            check(_lib.lib().qs_basis_states(self._h, _ptr(out)))
            self._states = out
        return self._states

    @property
    def states(self):
        return self._labels().copy()

    @property
    def reps(self):
        out = np.empty(self.Ns, dtype=np.uint64)
        check(_lib.lib().qs_basis_reps(self._h, _ptr(out)))
        return out

    def norms(self):
        """(orbit sizes, reference normalization at each label)."""
        orb = np.empty(self.Ns, dtype=np.uint32)
        ref = np.empty(self.Ns, dtype=np.float64)
        check(_lib.lib().qs_basis_norms(self._h, _ptr(orb), _ptr(ref)))
        return orb, ref

    def index(self, states):
        """Column of each state's orbit (-1 if outside the sector) and <state|column>."""
        s = np.ascontiguousarray(np.asarray(states, dtype=np.uint64))
        idx = np.empty(s.size, dtype=np.int64)
        amp = np.empty(s.size, dtype=np.complex128)
        check(_lib.lib().qs_basis_lookup(self._h, _ptr(s), s.size, _ptr(idx), _ptr(amp)))
        return idx, amp

    def _column(self, state):
        # _symmetry_column (discrete_basis.jl:2386-2395): the column holding the
        # state's orbit and the state's amplitude there
        if state < 0:
            raise ArgumentError("state must be nonnegative")
        idx, amp = self.index([state])
        if idx[0] < 0:
            raise ArgumentError("state has zero weight in this symmetry sector")
        return int(idx[0]), complex(amp[0])

    def state_index(self, state):
        """state_index(basis, s) -- basis.jl:1016-1027: position of s in states(basis)."""
        if state < 0:
            raise ArgumentError("state must be nonnegative")
        idx, _ = self.index([state])
        if idx[0] < 0 or int(self._labels()[idx[0]]) != int(state):
            raise ArgumentError("state %d is not represented by this basis" % int(state))
        return int(idx[0])

    def representative(self, state):
        """representative(basis, s) -- discrete_basis.jl:2397-2403: the label of s's column."""
        c, _ = self._column(state)
        return int(self._labels()[c])

    def get_amp(self, state):
        """get_amp(basis, s) -- discrete_basis.jl:2414-2422: s's amplitude in its column."""
        _, a = self._column(state)
        return a

    def normalization(self, state=None):
        """normalization(basis, s) -- discrete_basis.jl:2405-2412: inv(abs2(amp)); without a
        state, the normalization at every label (bit-exact with the reference's)."""
        orb, ref = self.norms()
        if state is None:
            return ref
        c, a = self._column(state)
        if int(self._labels()[c]) == int(state):
            return float(ref[c])
        return 1.0 / abs(a) ** 2

    def project_to(self, parent_vector, parent_states):
        """P' v for a vector given on explicit parent states (used for initial states)."""
        idx, amp = self.index(parent_states)
        out = np.zeros(self.Ns, dtype=np.complex128)
        ok = idx >= 0
        np.add.at(out, idx[ok], np.conj(amp[ok]) * np.asarray(parent_vector)[ok])
        return out


class SpinBasisGeneral(SpinBasis1D):
    """SpinBasisGeneral(L; nup, Nup, m, S, pauli, block_order, blocks...) -- basis.jl:530-620.

    Spin-1/2 states reduced by arbitrary site maps: each block is
    ``name=(map, quantum_number)``, the map one entry per site in QuSpin's
    zero-based labels (any entry zero or negative; ``-v`` inverts the site and
    sends it to site ``v``) or Julia's one-based labels (negative = inverted),
    the eigenvalue ``cis(2 pi q / order)`` with the map's order.  Stages run in
    ``block_order`` (default: the keyword order).  L <= 32: operators take
    two-bit hop couplings (the parent-table kernels).  More than 32 sites with
    blocks, or more than 63 sites (the reference's WideSpinBasis, basis.jl:
    361-445), up to 256: a wide basis (needs nup; states of more than 64
    sites are Python ints); operators take any monomial couplings.
    ``_explicit=True`` builds the wide form at any L.
    """

    def __init__(self, L, nup=None, Nup=None, m=None, S="1/2", pauli=True, block_order=None, _explicit=False,
                 **blocks):
        if str(S) not in ("1/2", "0.5"):
            raise ArgumentError("only spin-1/2 chains are on the GPU path")
        if nup is not None and Nup is not None:
            raise ArgumentError("specify only one of nup and Nup")
        sel = Nup if nup is None else nup
        if m is not None:
            if sel is not None:
                raise ArgumentError("m cannot be combined with nup or Nup")
            sel = int(round((float(m) + 0.5) * L))
        names = list(blocks) if block_order is None else [str(n) for n in block_order]
        for n in blocks:
            if n not in names:
                names.append(n)
        if any(n not in blocks for n in names):
            raise ArgumentError("block_order refers to an unknown symmetry block")
        maps, qn = [], []
        for n in names:
            spec = blocks[n]
            if not isinstance(spec, (tuple, list)) or len(spec) < 2:
                raise ArgumentError("general symmetry blocks must be (map, quantum_number)")
            mp = [int(v) for v in spec[0]]
            if len(mp) != int(L):
                raise ArgumentError("a general symmetry map must contain one entry per site")
            maps.append(mp)
            qn.append(int(spec[1]))
        self.L = int(L)
        self.nup = None if sel is None else int(sel)
        self.pauli = bool(pauli)
        self.a = 1
        self._blocks = dict(kblock=None, pblock=None, zblock=None, pzblock=None, zAblock=None, zBblock=None)
        self._general = [(n, tuple(mp), q) for n, mp, q in zip(names, maps, qn)]
        mats = np.ascontiguousarray(np.asarray(maps or [[0]], dtype=np.int32))
        qs = np.ascontiguousarray(np.asarray(qn or [0], dtype=np.int32))
        h = ctypes.c_void_p()
        flags = _lib.QS_BASIS_EXPLICIT if _explicit else 0
        self.wide = bool(_explicit) or self.L > 63 or (self.L > 32 and len(maps) > 0)
        check(_lib.lib().qs_basis_create_general(self.L, _blk(self.nup), int(self.pauli), len(maps), _ptr(mats),
                                                  _ptr(qs), flags, ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        check(_lib.lib().qs_basis_size(self._h, ctypes.byref(n)))
        self.Ns = int(n.value)
        c = ctypes.c_int()
        check(_lib.lib().qs_basis_is_complex(self._h, ctypes.byref(c)))
        self.is_complex = bool(c.value)
        self._states = None

    def state_words(self, which=0):
        """(Ns, 4) uint64 words of states(basis) (which=0) or of the orbit
        minima (which=1); site i is bit i % 64 of word i // 64."""
        out = np.empty((self.Ns, 4), dtype=np.uint64)
        check(_lib.lib().qs_basis_states_wide(self._h, int(which), _ptr(out)))
        return out

    def _labels(self):
        if self.L <= 64:
            return SpinBasis1D._labels(self)
        if self._states is None:
            w = self.state_words(0)
            self._states = np.array([int(r[0]) | int(r[1]) << 64 | int(r[2]) << 128 | int(r[3]) << 192 for r in w],
                                    dtype=object)
        return self._states

    def index(self, states):
        """Column of each state's orbit (-1 if outside the sector) and <state|column>;
        states of more than 64 sites as Python ints."""
        if self.L <= 64:
            return SpinBasis1D.index(self, states)
        vals = [int(v) for v in np.atleast_1d(np.asarray(states, dtype=object))]
        words = np.array([[(v >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(4)] for v in vals],
                         dtype=np.uint64).reshape(-1, 4)
        words = np.ascontiguousarray(words)
        idx = np.empty(len(vals), dtype=np.int64)
        amp = np.empty(len(vals), dtype=np.complex128)
        check(_lib.lib().qs_basis_lookup_wide(self._h, _ptr(words), len(vals), _ptr(idx), _ptr(amp)))
        return idx, amp

    @property
    def reps(self):
        if self.L <= 64:
            return SpinBasis1D.reps.fget(self)
        w = self.state_words(1)
        return np.array([int(r[0]) | int(r[1]) << 64 | int(r[2]) << 128 | int(r[3]) << 192 for r in w], dtype=object)

    @property
    def blocks(self):
        out = {"nup": self.nup, "pauli": self.pauli, "block_order": [n for n, _, _ in self._general]}
        for n, mp, q in self._general:
            out[n] = q
        return out
