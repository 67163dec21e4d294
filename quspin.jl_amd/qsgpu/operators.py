"""QuantumLinearOperator and its Krylov consumers on the GPU.

Mirrors operators.jl: OperatorTerm (:42-58), QuantumLinearOperator
(:2331-2346, :2492-2552), mul! (:3147-3208), transpose/conj/adjoint
(:3286-3296), eigsh (:3367-3408), evolve for a static generator
(:2073-2192) and ExpmMultiplyParallel (tools.jl:1605-1705).
"""
import ctypes
import numbers

import numpy as np

from . import _lib
from .basis import SpinBasis1D
from .errors import ArgumentError, DimensionMismatch, check

_DT = {np.dtype(np.float64): _lib.QS_F64, np.dtype(np.complex128): _lib.QS_C128,
       np.dtype(np.float32): _lib.QS_F32, np.dtype(np.complex64): _lib.QS_C64}


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_info():
    name = ctypes.create_string_buffer(256)
    hbm = ctypes.c_int64()
    ncu = ctypes.c_int()
    check(_lib.lib().qs_device_info(name, 256, ctypes.byref(hbm), ctypes.byref(ncu)))
    return {"name": name.value.decode(), "hbm_bytes": hbm.value, "n_cu": ncu.value}


class OperatorTerm:
    """A local spin operator string with couplings (coefficient, sites...)."""

    def __init__(self, op, couplings):
        if not isinstance(op, str) or len(op) == 0:
            raise ArgumentError("operator string cannot be empty")
        arity = sum(1 for c in op if c != "|")
        normalized = []
        for c in couplings:
            c = tuple(c)
            if len(c) != arity + 1:
                raise ArgumentError("operator arity and coupling sites differ")
            if not all(isinstance(s, numbers.Integral) for s in c[1:]):
                raise ArgumentError("all sites must be integers")
            normalized.append(c)
        self.op = op
        self.couplings = normalized

    def __repr__(self):
        return "OperatorTerm(%r, %r)" % (self.op, self.couplings)

    def __eq__(self, other):
        return isinstance(other, OperatorTerm) and self.op == other.op and self.couplings == other.couplings


def _flatten(terms):
    ops, ncoup, coef, sites = [], [], [], []
    complex_type = False
    for t in terms:
        if isinstance(t, (tuple, list)):
            t = OperatorTerm(t[0], t[1])
        ops.append(t.op.encode())
        ncoup.append(len(t.couplings))
        for c in t.couplings:
            if isinstance(c[0], complex) or np.iscomplexobj(np.asarray(c[0])):
                complex_type = True
            z = complex(c[0])
            coef += [z.real, z.imag]
            sites += [int(s) for s in c[1:]]
    arr = (ctypes.c_char_p * max(1, len(ops)))(*ops)
    return (arr, np.asarray(ncoup, dtype=np.int32), np.asarray(coef if coef else [0.0], dtype=np.float64),
            np.asarray(sites if sites else [0], dtype=np.int32), complex_type, len(ops))


class QuantumLinearOperator:
    """Matrix-free sector operator P'·H·P applied by the gfx950 kernels."""

    def __init__(self, basis, static_list, diagonal=None, check_symm=True, check_herm=True, check_pcon=True,
                 _transposed=False, _conjugated=False, _shared=None):
        self.basis = basis
        self.static_list = [t if isinstance(t, OperatorTerm) else OperatorTerm(t[0], t[1]) for t in static_list]
        self.transposed = _transposed
        self.conjugated = _conjugated
        if _shared is not None:
            self._shared = _shared
            return
        arr, ncoup, coef, sites, complex_type, nterm = _flatten(self.static_list)
        flags = (_lib.QS_CHECK_SYMM if check_symm else 0) | (_lib.QS_CHECK_HERM if check_herm else 0) | \
            (_lib.QS_CHECK_PCON if check_pcon else 0)
        d = None
        if diagonal is not None:
            diagonal = np.asarray(diagonal)
            if diagonal.shape != (basis.Ns,):
                raise DimensionMismatch("diagonal must have length Ns")
            if np.iscomplexobj(diagonal):
                flags |= _lib.QS_FORCE_COMPLEX
            dd = np.ascontiguousarray(diagonal.astype(np.complex128))
            d = _ptr(dd)
        if complex_type:
            flags |= _lib.QS_FORCE_COMPLEX
        h = ctypes.c_void_p()
        check(_lib.lib().qs_op_create(basis._h, nterm, arr, _ptr(ncoup), _ptr(coef), _ptr(sites), d, flags,
                                       ctypes.byref(h)))
        self._shared = _Handle(h)
        self._diag = None if diagonal is None else np.array(diagonal)

    # --------------------------------------------------------------- protocol
    @property
    def _h(self):
        return self._shared.h

    def _info(self):
        c, hm, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        check(_lib.lib().qs_op_info(self._h, ctypes.byref(c), ctypes.byref(hm), ctypes.byref(n)))
        return bool(c.value), bool(hm.value), int(n.value)

    @property
    def dtype(self):
        return np.dtype(np.complex128) if self._info()[0] else np.dtype(np.float64)

    @property
    def hermitian(self):
        return self._info()[1]

    def ishermitian(self):
        return self.hermitian

    @property
    def Ns(self):
        return self.basis.Ns

    @property
    def shape(self):
        return (self.Ns, self.Ns)

    @property
    def diagonal(self):
        return np.zeros(self.Ns, dtype=self.dtype) if self._diag is None else self._diag.copy()

    def set_diagonal(self, diagonal):
        diagonal = np.asarray(diagonal)
        if diagonal.shape != (self.Ns,):
            raise DimensionMismatch("diagonal must have length Ns")
        dd = np.ascontiguousarray(diagonal.astype(np.complex128))
        check(_lib.lib().qs_op_set_diagonal(self._h, _ptr(dd)))
        self._diag = diagonal.copy()
        return self

    def _flagged(self, t, c):
        return QuantumLinearOperator(self.basis, self.static_list, _transposed=self.transposed ^ t,
                                     _conjugated=self.conjugated ^ c, _shared=self._shared)

    def transpose(self):
        return self._flagged(True, False)

    def conj(self):
        return self._flagged(False, True)

    def adjoint(self):
        return self._flagged(True, True)

    @property
    def T(self):
        return self.transpose()

    @property
    def H(self):
        return self.adjoint()

    # ------------------------------------------------------------- products
    def _vec_dtype(self, x_dtype, prefer32=False):
        cplx = self._info()[0] or np.issubdtype(x_dtype, np.complexfloating)
        if prefer32:
            return np.dtype(np.complex64) if cplx else np.dtype(np.float32)
        return np.dtype(np.complex128) if cplx else np.dtype(np.float64)

    def mul(self, out, x, alpha=1.0, beta=0.0):
        """mul!(out, op, x, alpha, beta): out = beta*out + alpha*op*x (vector or matrix columns)."""
        x = np.asarray(x)
        if x.shape[0] != self.Ns:
            raise DimensionMismatch("operator and value dimensions do not match")
        if out.shape != x.shape:
            raise DimensionMismatch("output and operator dimensions do not match")
        dt = out.dtype
        if dt not in _DT:
            raise ArgumentError("unsupported output dtype %s" % dt)
        ncols = 1 if x.ndim == 1 else x.shape[1]
        xx = np.asfortranarray(x.astype(dt, copy=False)) if x.ndim == 2 else np.ascontiguousarray(x.astype(dt, copy=False))
        yy = np.asfortranarray(out) if out.ndim == 2 else np.ascontiguousarray(out)
        a, b = complex(alpha), complex(beta)
        check(_lib.lib().qs_op_matvec(self._h, _DT[dt], _ptr(xx), _ptr(yy), ncols, a.real, a.imag, b.real, b.imag,
                                       int(self.transposed), int(self.conjugated), 0))
        if yy is not out:
            out[...] = yy
        return out

    def matvec(self, x, dtype=None):
        x = np.asarray(x)
        dt = np.dtype(dtype) if dtype is not None else self._vec_dtype(x.dtype)
        out = np.zeros(x.shape, dtype=dt)
        return self.mul(out, x)

    def __matmul__(self, x):
        return self.matvec(x)

    def __mul__(self, x):
        return self.matvec(x)

    def dot(self, x):
        return self.matvec(x)

    def matvec_device(self, x, y, alpha=1.0, beta=0.0):
        """y = beta*y + alpha*op*x on device tensors (torch, HBM-resident)."""
        import torch
        tmap = {torch.float64: _lib.QS_F64, torch.complex128: _lib.QS_C128, torch.float32: _lib.QS_F32,
                torch.complex64: _lib.QS_C64}
        if x.dtype != y.dtype or x.dtype not in tmap:
            raise ArgumentError("x and y must share a supported dtype")
        if x.shape[0] != self.Ns or y.shape != x.shape:
            raise DimensionMismatch("operator and value dimensions do not match")
        ncols = 1 if x.dim() == 1 else x.shape[1]
        if x.dim() == 2 and not (x.t().is_contiguous() and y.t().is_contiguous()):
            raise ArgumentError("matrix arguments must be column-major (use .t().contiguous().t())")
        a, b = complex(alpha), complex(beta)
        check(_lib.lib().qs_op_matvec(self._h, tmap[x.dtype], ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(y.data_ptr()), ncols, a.real, a.imag, b.real, b.imag,
                                       int(self.transposed), int(self.conjugated), 1))
        return y

    def set_variant(self, lanes_per_row):
        """Fix the matvec kernel variant: 0 = automatic, 1/2/4/8 = the general kernel
        with that many lanes per row, VARIANT_HOP (+1, +2, +3) = the two-bit hop kernel,
        VARIANT_SPLIT = the split kernel (hop couplings, no symmetry or spin inversion only),
        VARIANT_SELL = plain products from the assembled matrix (assembled by the first product)."""
        check(_lib.lib().qs_op_set_variant(self._h, int(lanes_per_row)))
        return self

    def set_graphs(self, enable=True):
        """Replay the Krylov exponential's halving recursion from a HIP graph (on a
        non-blocking stream set with set_stream; no legacy-default-stream work in
        other threads meanwhile)."""
        check(_lib.lib().qs_op_set_graphs(self._h, int(bool(enable))))
        return self

    def assemble(self):
        """Assemble the sector matrix on the device (the reference's
        Hamiltonian(...; static_fmt=:csr), operators.jl:461-576); plain products then
        read the stored rows.  Needs two-bit hop couplings."""
        check(_lib.lib().qs_op_assemble(self._h))
        self.assembled = True
        return self

    def tocsr(self):
        """The assembled matrix as scipy.sparse.csr_matrix (duplicates merged,
        exact zeros dropped; SparseMatrixCSR, sparse_formats.jl:21-50)."""
        import scipy.sparse as sp
        L = _lib.lib()
        nnz = ctypes.c_int64(0)
        check(L.qs_op_csr(self._h, ctypes.byref(nnz), None, None, None))
        n = self.Ns
        rowptr = np.empty(n + 1, dtype=np.int64)
        col = np.empty(max(1, nnz.value), dtype=np.int64)
        val = np.empty(2 * max(1, nnz.value), dtype=np.float64)
        check(L.qs_op_csr(self._h, ctypes.byref(nnz), rowptr.ctypes.data, col.ctypes.data, val.ctypes.data))
        v = val[:2 * nnz.value].view(np.complex128)
        if not np.any(v.imag):
            v = v.real.copy()
        return sp.csr_matrix((v, col[:nnz.value], rowptr), shape=(n, n))

    def tune(self, dtype=None, repeats=5):
        """Time every kernel variant for `dtype` and keep the fastest; returns (lanes_per_row, ms)."""
        dt = np.dtype(dtype) if dtype is not None else self._vec_dtype(np.float64)
        lpr, ms = ctypes.c_int(), ctypes.c_double()
        check(_lib.lib().qs_op_tune(self._h, _DT[dt], int(repeats), ctypes.byref(lpr), ctypes.byref(ms)))
        return lpr.value, ms.value

    def set_stream(self, stream_ptr):
        check(_lib.lib().qs_op_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def toarray(self):
        """Matrix(op): apply to the identity (small sectors only)."""
        eye = np.eye(self.Ns, dtype=self._vec_dtype(np.float64))
        out = np.zeros_like(eye)
        return self.mul(out, eye)


class _Handle:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        if self.h is not None and self.h.value:
            try:
                _lib.lib().qs_op_destroy(self.h)
            except (TypeError, AttributeError):
                pass   # interpreter shutdown: module globals are already gone
            self.h = None


def apply(op, x, out=None, a=1.0):
    """apply(op, x; out, a) -- operators.jl:3311-3322."""
    if out is None:
        out = np.zeros(np.asarray(x).shape, dtype=op._vec_dtype(np.asarray(x).dtype))
    return op.mul(out, x, alpha=a, beta=0.0)


def mul(out, op, x, alpha=1.0, beta=0.0):
    """mul!(out, op, x, alpha, beta)."""
    return op.mul(out, x, alpha, beta)


def eigsh(op, k=None, which="SA", tol=0.0, maxiter=300, ncv=None, v0=None, return_eigenvectors=True,
          dtype=None):
    """eigsh(op; k, which, tol, maxiter, ncv, v0, return_eigenvectors) -- operators.jl:3367-3408.

    ``dtype`` selects the Krylov storage: float64/complex128 (FP64 form) or
    float32/complex64 (FP32-stream form, FP64 reductions)."""
    n = op.Ns
    if k is None:
        k = min(6, n)
    if not (1 <= k <= n):
        raise ArgumentError("k must lie in 1:Ns")
    w = str(which).upper()
    if w not in _lib.WHICH:
        raise ArgumentError("which must be SA, LA, SM, LM, or BE")
    cplx, herm, _ = op._info()
    # a non-Hermitian operator takes Arpack.eigs' complex path (operators.jl:3383-3390)
    cplx = cplx or not herm
    if dtype is None:
        dt = np.dtype(np.complex128) if cplx else np.dtype(np.float64)
    else:
        dt = np.dtype(dtype)
        if cplx and not np.issubdtype(dt, np.complexfloating):
            dt = np.dtype(np.complex64) if dt == np.float32 else np.dtype(np.complex128)
    vals = np.zeros(k if herm else 2 * k, dtype=np.float64)
    vecs = np.zeros((n, k), dtype=dt, order="F") if return_eigenvectors else None
    v0p = None
    if v0 is not None:
        v0a = np.ascontiguousarray(np.asarray(v0).astype(dt))
        if v0a.shape != (n,):
            raise DimensionMismatch("v0 must have length Ns")
        v0p = _ptr(v0a)
    nconv, nmv = ctypes.c_int(), ctypes.c_int()
    st = _lib.lib().qs_eigsh(op._h, _DT[dt], int(k), _lib.WHICH[w], float(tol), int(maxiter),
                              int(ncv) if ncv else 0, v0p, _ptr(vals), _ptr(vecs) if vecs is not None else None,
                              ctypes.byref(nconv), ctypes.byref(nmv))
    check(st)
    eigsh.last_n_matvec = nmv.value
    if not herm:
        vals = vals[0::2] + 1j * vals[1::2]
    if return_eigenvectors:
        return vals, vecs
    return vals


def expm_multiply(op, v, scales, tol=1e-12, krylov_dim=30, dtype=np.complex128):
    """exp(scale*op) v for each scale -- _krylov_expmv_times (operators.jl:1805-1893)."""
    dt = np.dtype(dtype)
    if not np.issubdtype(dt, np.complexfloating):
        raise ArgumentError("expm_multiply needs a complex dtype")
    v = np.ascontiguousarray(np.asarray(v).astype(dt))
    if v.shape != (op.Ns,):
        raise DimensionMismatch("operator and vector dimensions do not match")
    sc = np.asarray(scales, dtype=np.complex128).ravel()
    out = np.zeros((op.Ns, sc.size), dtype=dt, order="F")
    nmv = ctypes.c_int()
    scv = np.ascontiguousarray(sc)
    check(_lib.lib().qs_expm_multiply(op._h, _DT[dt], _ptr(v), sc.size, _ptr(scv), float(tol), int(krylov_dim),
                                       _ptr(out), ctypes.byref(nmv)))
    expm_multiply.last_n_matvec = nmv.value
    return out


class Hamiltonian:
    """Hamiltonian(static_list, dynamic_list; basis, static_fmt, check_*) -- operators.jl:578-614, 648-737:
    the static terms form one QuantumLinearOperator on `basis` and each group of dynamic entries
    ``[op, couplings, f, f_args]`` sharing ``(f, f_args)`` another, whose product is scaled by
    ``f(time, *f_args)``.  ``static_fmt`` "csr" / "csc" assembles the static sector matrix on the
    device (its products then stream the stored rows); None / "dense" keeps it matrix-free."""

    def __init__(self, static_list, dynamic_list=(), basis=None, static_fmt=None, check_symm=True, check_herm=True,
                 check_pcon=True):
        if basis is None:
            raise ArgumentError("a basis is required")
        fmt = None if static_fmt is None else str(static_fmt).lower().lstrip(":")
        if fmt not in (None, "dense", "csr", "csc"):
            raise ArgumentError("static_fmt must be :dense, :csr, or :csc")
        self.basis = basis
        kw = dict(check_symm=check_symm, check_herm=check_herm, check_pcon=check_pcon)
        self.static = QuantumLinearOperator(basis, list(static_list), **kw)
        if fmt in ("csr", "csc"):
            self.static.assemble()
        # entries sharing (f, f_args) form one operator, checked as one
        # collection (_dynamic_groups, discrete_basis.jl:2260-2309)
        groups = []
        for entry in dynamic_list:
            if len(entry) != 4:
                raise ArgumentError("dynamic operator entries are [op, couplings, f, f_args]")
            op, couplings, f, args = entry
            key = (f, tuple(args))
            for g in groups:
                if g[0][0] is key[0] and g[0][1] == key[1]:
                    g[1].append((op, couplings))
                    break
            else:
                groups.append((key, [(op, couplings)]))
        self.dynamic = [(QuantumLinearOperator(basis, terms, **kw), key[0], key[1]) for key, terms in groups]

    @property
    def Ns(self):
        return self.basis.Ns

    @property
    def shape(self):
        return (self.Ns, self.Ns)

    def apply(self, v, time=0.0):
        """H(time) v (operators.jl:1383-1420)."""
        out = np.asarray(self.static.matvec(np.asarray(v, dtype=np.complex128)), dtype=np.complex128)
        for op, f, args in self.dynamic:
            c = f(time, *args)
            if c == 0:
                continue
            out = out + c * op.matvec(np.asarray(v, dtype=np.complex128))
        return out


_DRIVE = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def _evolve_dynamic(H, v0, t0, targets, imag_time, max_step, rtol, atol, dtype):
    # operators.jl:2092-2128 on the device (qs_evolve_dynamic)
    dt = np.dtype(dtype)
    if not np.issubdtype(dt, np.complexfloating):
        raise ArgumentError("evolve needs a complex dtype")
    v = np.asarray(v0)
    if v.shape[0] != H.Ns:
        raise DimensionMismatch("Hamiltonian and value dimensions do not match")
    ncols = 1 if v.ndim == 1 else v.shape[1]
    vv = np.asfortranarray(v.astype(dt).reshape(H.Ns, ncols))
    errors = []

    def drive(t, n, coef, _user):
        try:
            for j, (_, f, args) in enumerate(H.dynamic):
                c = complex(f(t, *args))
                coef[2 * j] = c.real
                coef[2 * j + 1] = c.imag
        except Exception as e:   # noqa: BLE001 -- reported after the call
            errors.append(e)
            for j in range(n):
                coef[2 * j] = coef[2 * j + 1] = 0.0

    cb = _DRIVE(drive)
    ops = (ctypes.c_void_p * max(1, len(H.dynamic)))(*[op._h for op, _, _ in H.dynamic])
    tt = np.ascontiguousarray(targets, dtype=np.float64)
    out = np.zeros((H.Ns, ncols, tt.size), dtype=dt, order="F")
    steps, nmv = ctypes.c_int64(), ctypes.c_int64()
    st = _lib.lib().qs_evolve_dynamic(H.static._h, len(H.dynamic), ops, cb, None, _DT[dt], _ptr(vv), ncols,
                                       float(t0), tt.size, _ptr(tt), int(bool(imag_time)), float(rtol), float(atol),
                                       float(max_step), _ptr(out), ctypes.byref(steps), ctypes.byref(nmv))
    if errors:
        raise errors[0]
    check(st)
    evolve.last_n_steps = steps.value
    evolve.last_n_matvec = nmv.value
    return out if v.ndim == 2 else out[:, 0, :]


def evolve(op, v0, t0, times, imag_time=False, iterate=False, tol=None, krylov_dim=30, dtype=np.complex128,
           max_step=0.01, rtol=None, atol=1e-11):
    """evolve(H, v0, t0, times) -- operators.jl:2073-2192 (:SE branch).  A static generator
    (QuantumLinearOperator, or a Hamiltonian without dynamic terms) takes the Krylov exponential;
    a Hamiltonian with dynamic terms the adaptive RK45 integrator (max_step, rtol = tol or 1e-9,
    atol)."""
    scalar = isinstance(times, numbers.Number)
    if isinstance(op, Hamiltonian):
        if op.dynamic:
            if not max_step > 0:
                raise ArgumentError("max_step must be positive")
            targets = np.atleast_1d(np.asarray(times, dtype=np.float64))
            r = rtol if rtol is not None else (tol if tol is not None else 1e-9)
            states = _evolve_dynamic(op, v0, t0, targets, imag_time, max_step, r, atol, dtype)
            if iterate:
                return (states[..., j].copy() for j in range(states.shape[-1]))
            if scalar:
                return states[..., 0].copy()
            return states
        op = op.static
    if tol is None:
        tol = 1e-12
    targets = np.atleast_1d(np.asarray(times, dtype=np.float64))
    offsets = targets - float(t0)
    scales = -offsets.astype(np.complex128) if imag_time else -1j * offsets
    states = expm_multiply(op, v0, scales, tol=tol, krylov_dim=krylov_dim, dtype=dtype)
    if imag_time:
        states /= np.linalg.norm(states, axis=0, keepdims=True)
    if iterate:
        return (states[:, j].copy() for j in range(states.shape[1]))
    if scalar:
        return states[:, 0].copy()
    return states


class ExpmMultiplyParallel:
    """exp(a*A) v via the Krylov action (tools.jl:1605-1705)."""

    def __init__(self, A, a=1.0, dtype=None, copy=False):
        if A.shape[0] != A.shape[1]:
            raise ArgumentError("A must be a square matrix")
        self.A = A
        self.a = complex(a)
        self.dtype = np.dtype(np.complex128 if dtype is None else dtype)

    def set_a(self, a, dtype=None):
        if dtype is not None and np.dtype(dtype) != self.dtype:
            raise ArgumentError("changing the exponential dtype requires constructing a new operator")
        self.a = complex(a)
        return self

    def apply(self, v, work_array=None, overwrite_v=False, tol=None):
        v = np.asarray(v)
        if v.shape[0] != self.A.shape[1]:
            raise DimensionMismatch("A and v dimensions do not match")
        if work_array is not None and np.asarray(work_array).size != 2 * v.size:
            raise DimensionMismatch("work_array must contain twice as many elements as v")
        t = 1e-12 if tol is None else tol
        if v.ndim == 1:
            res = expm_multiply(self.A, v, [self.a], tol=t, dtype=self.dtype)[:, 0]
        else:
            res = np.stack([expm_multiply(self.A, v[:, j], [self.a], tol=t, dtype=self.dtype)[:, 0]
                            for j in range(v.shape[1])], axis=1)
        if overwrite_v:
            v[...] = res
            return v
        return res

    def __mul__(self, v):
        return self.apply(v)

    __matmul__ = __mul__
