"""Serving loop: eigsh and evolve requests on one GPU (qs_server_* of libqsgpu).

Every request runs exactly the reference's solver (eigsh, operators.jl:3367-3408;
evolve / _krylov_expmv_times, operators.jl:2073-2192, 1805-1893) on the GPU.
The native server admits requests first-in first-out as workers and HBM
budget free up, runs each on its worker's own stream so their kernels
overlap, and reports per-request queue / run / total seconds.
"""
import ctypes
import numbers

import numpy as np

from . import _lib
from .errors import ArgumentError, check
from .operators import _DT, QuantumLinearOperator

QS_JOB_EIGSH, QS_JOB_EXPM = 0, 1


class _Job(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("op", ctypes.c_void_p), ("dtype", ctypes.c_int), ("k", ctypes.c_int),
                ("which", ctypes.c_int), ("maxiter", ctypes.c_int), ("ncv", ctypes.c_int), ("tol", ctypes.c_double),
                ("v0", ctypes.c_void_p), ("n_times", ctypes.c_int64), ("scales", ctypes.c_void_p),
                ("krylov_dim", ctypes.c_int), ("want_vectors", ctypes.c_int)]


def _bind():
    L = _lib.lib()
    if not getattr(L, "_qs_serve_bound", False):
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.qs_server_create.argtypes = [i32, i64, ctypes.POINTER(ctypes.c_void_p)]
        L.qs_server_destroy.argtypes = [vp]
        L.qs_server_submit.argtypes = [vp, ctypes.POINTER(_Job), ctypes.POINTER(ctypes.c_int64)]
        L.qs_server_wait.argtypes = [vp, i64, vp, vp, vp, vp, vp]
        for n in ("qs_server_create", "qs_server_destroy", "qs_server_submit", "qs_server_wait"):
            getattr(L, n).restype = ctypes.c_int
        L._qs_serve_bound = True
    return L


class Server:
    """Request loop with `n_workers` overlapping workers and an HBM budget (bytes)."""

    def __init__(self, n_workers=4, hbm_budget=200e9):
        h = ctypes.c_void_p()
        check(_bind().qs_server_create(int(n_workers), int(hbm_budget), ctypes.byref(h)))
        self._h = h
        self._pending = {}

    def close(self):
        if self._h is not None and self._h.value:
            check(_bind().qs_server_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except (TypeError, AttributeError):
            pass   # interpreter shutdown: module globals are already gone

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _submit(self, job, keep):
        t = ctypes.c_int64()
        check(_bind().qs_server_submit(self._h, ctypes.byref(job), ctypes.byref(t)))
        self._pending[t.value] = keep
        return t.value

    def submit_eigsh(self, op, k=1, which="SA", tol=1e-10, v0=None, maxiter=300, ncv=None, dtype=None,
                     return_eigenvectors=False):
        """Queue eigsh(op; k, which, tol, v0, maxiter, ncv); returns a ticket."""
        w = str(which).upper()
        if w not in _lib.WHICH:
            raise ArgumentError("which must be SA, LA, SM, LM, or BE")
        cplx, herm, _ = op._info()
        cplx = cplx or not herm   # non-Hermitian: complex Arnoldi
        dt = np.dtype(dtype) if dtype is not None else np.dtype(np.complex128 if cplx else np.float64)
        if cplx and not np.issubdtype(dt, np.complexfloating):
            dt = np.dtype(np.complex64) if dt == np.float32 else np.dtype(np.complex128)
        v = None if v0 is None else np.ascontiguousarray(np.asarray(v0).astype(dt))
        job = _Job(QS_JOB_EIGSH, op._h, _DT[dt], int(k), _lib.WHICH[w], int(maxiter), int(ncv or 0), float(tol),
                   None if v is None else v.ctypes.data, 0, None, 0, int(bool(return_eigenvectors)))
        return self._submit(job, dict(kind="eigsh", op=op, k=int(k), dt=dt, herm=op._info()[1],
                                      vecs=bool(return_eigenvectors)))

    def submit_evolve(self, op, v0, t0, times, imag_time=False, tol=1e-12, krylov_dim=30, dtype=np.complex128):
        """Queue evolve(op, v0, t0, times) of a static generator; returns a ticket."""
        dt = np.dtype(dtype)
        if not np.issubdtype(dt, np.complexfloating):
            raise ArgumentError("evolve needs a complex dtype")
        targets = np.atleast_1d(np.asarray(times, dtype=np.float64))
        off = targets - float(t0)
        sc = np.ascontiguousarray((-off.astype(np.complex128)) if imag_time else (-1j * off))
        v = np.ascontiguousarray(np.asarray(v0).astype(dt))
        job = _Job(QS_JOB_EXPM, op._h, _DT[dt], 0, 0, 0, 0, float(tol), v.ctypes.data, sc.size, sc.ctypes.data,
                   int(krylov_dim), 0)
        return self._submit(job, dict(kind="evolve", op=op, nt=sc.size, dt=dt, imag=bool(imag_time),
                                      scalar=isinstance(times, numbers.Number)))

    def wait(self, ticket):
        """Results of one request: {'vals' | 'states', 'seconds': (queued, running, total), 'n_matvec'}."""
        keep = self._pending.pop(ticket)
        op = keep["op"]
        n = op.Ns
        sec = np.zeros(3)
        nconv, nmv = ctypes.c_int(), ctypes.c_int()
        if keep["kind"] == "eigsh":
            k = keep["k"]
            vals = np.zeros(k if keep["herm"] else 2 * k)
            out = np.zeros((n, k), dtype=keep["dt"], order="F") if keep["vecs"] else None
            check(_bind().qs_server_wait(self._h, ticket, vals.ctypes.data, None if out is None else out.ctypes.data,
                                         ctypes.byref(nconv), ctypes.byref(nmv), sec.ctypes.data))
            if not keep["herm"]:
                vals = vals[0::2] + 1j * vals[1::2]
            res = dict(vals=vals, vecs=out)
        else:
            out = np.zeros((n, max(1, keep["nt"])), dtype=keep["dt"], order="F")
            check(_bind().qs_server_wait(self._h, ticket, None, out.ctypes.data, ctypes.byref(nconv),
                                         ctypes.byref(nmv), sec.ctypes.data))
            if keep["imag"]:
                out /= np.linalg.norm(out, axis=0, keepdims=True)
            res = dict(states=out[:, 0].copy() if keep["scalar"] else out)
        res["seconds"] = tuple(sec)
        res["n_matvec"] = nmv.value
        return res


def poisson_workload(n_requests, rate=200.0, seed=0, Ls=(16, 20, 24, 28)):
    """The mixed request stream of the serving benchmark: Poisson arrivals (mean `rate` per s),
    60% eigsh k in 1..4 and 40% evolve to t ~ U[1, 10]; XXZ / J1-J2 chains with L drawn from Ls,
    nup = L/2, a random symmetry sector, J2 ~ U[0, 0.5], Delta ~ U[0, 2]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    for i in range(n_requests):
        L = int(rng.choice(Ls))
        k = int(rng.integers(0, L))
        blocks = dict(kblock=k)
        if k in (0, L // 2):
            p = rng.choice([None, 1, -1])
            if p is not None:
                blocks["pblock"] = int(p)
        z = rng.choice([None, 1, -1])
        if z is not None:
            blocks["zblock"] = int(z)
        req = dict(arrival=float(t[i]), L=L, nup=L // 2, blocks=blocks, J2=float(rng.uniform(0, 0.5)),
                   delta=float(rng.uniform(0, 2)), seed=int(rng.integers(0, 2**31)))
        if rng.uniform() < 0.6:
            req.update(kind="eigsh", k=int(rng.integers(1, 5)))
        else:
            req.update(kind="evolve", t=float(rng.uniform(1, 10)))
        reqs.append(req)
    return reqs


def chain_terms(L, J2, delta, J=1.0):
    """XXZ / J1-J2 chain couplings (periodic), pauli=False conventions."""
    b1 = [(i, i % L + 1) for i in range(1, L + 1)]
    terms = [("+-", [(J / 2, i, j) for i, j in b1]), ("-+", [(J / 2, i, j) for i, j in b1]),
             ("zz", [(J * delta, i, j) for i, j in b1])]
    if J2:
        b2 = [(i, (i + 1) % L + 1) for i in range(1, L + 1)]
        terms += [("+-", [(J2 / 2, i, j) for i, j in b2]), ("-+", [(J2 / 2, i, j) for i, j in b2]),
                  ("zz", [(J2 * delta, i, j) for i, j in b2])]
    return terms


class RequestBuilder:
    """Bases cached per sector; one operator per request (its couplings are the request's).
    With `assemble`, sectors of at most `assemble_rows` rows are assembled on the device
    (the reference's static_fmt path), so the request's products read stored rows."""

    def __init__(self, assemble=False, assemble_rows=4_000_000):
        from .basis import SpinBasis1D
        self._SB = SpinBasis1D
        self.bases = {}
        self.assemble = bool(assemble)
        self.assemble_rows = int(assemble_rows)

    def operator(self, req):
        key = (req["L"], tuple(sorted(req["blocks"].items())))
        if key not in self.bases:
            self.bases[key] = self._SB(req["L"], nup=req["nup"], pauli=False, **req["blocks"])
        b = self.bases[key]
        op = QuantumLinearOperator(b, chain_terms(req["L"], req["J2"], req["delta"]))
        if self.assemble and op.Ns <= self.assemble_rows:
            try:
                # assembled by the first product, on the worker stream that runs it
                op.set_variant(_lib.VARIANT_SELL)
            except ArgumentError:
                pass
        return op

    def initial_state(self, req, op):
        rng = np.random.default_rng(req["seed"])
        v = rng.standard_normal(op.Ns) + 0j
        return v / np.linalg.norm(v)
