"""Lanczos tools on the GPU operator: lanczos_full, lanczos_iter, lin_comb_Q_T,
expm_lanczos -- tools.jl:2742-2952.

The recurrence runs on the device through qs_lanczos (the reference's
three-term Lanczos, step for step, with optional sequential full
orthogonalisation); the small SymTridiagonal eigenproblem is solved in FP64
on the host.  The operator must be a QuantumLinearOperator: the GPU path has
no generic-matrix fallback.
"""
import ctypes

import numpy as np

from . import _lib
from .errors import ArgumentError, DimensionMismatch, check
from .operators import _DT, QuantumLinearOperator


def _operator(A):
    if not isinstance(A, QuantumLinearOperator):
        raise ArgumentError("the GPU Lanczos tools act on a QuantumLinearOperator")
    if A.transposed or A.conjugated:
        raise ArgumentError("the GPU Lanczos tools act on the operator itself (no transpose / conj view)")
    return A


def _run(A, v0, m, full_ortho, eps, keep_rows, out=None):
    op = _operator(A)
    v0 = np.asarray(v0)
    n = op.Ns
    if v0.shape != (n,):
        raise DimensionMismatch("operator and vector dimensions do not match")
    m = int(m)
    if not (1 <= m < n):
        raise ArgumentError("Lanczos dimension m must satisfy 1 <= m < length(v0)")
    dt = op._vec_dtype(v0.dtype, prefer32=v0.dtype in (np.float32, np.complex64))
    v = np.ascontiguousarray(v0.astype(dt))
    alphas, betas = np.zeros(m), np.zeros(max(1, m - 1))
    E, V = np.zeros(m), np.zeros(m * m)
    Q = None
    if keep_rows:
        if out is not None:
            if out.shape != (m, n) or out.dtype != dt or not out.flags.c_contiguous:
                raise DimensionMismatch("out must have shape (m, length(v0))")
            Q = out
        else:
            Q = np.zeros((m, n), dtype=dt)
    used = ctypes.c_int()
    check(_lib.lib().qs_lanczos(op._h, _DT[dt], v.ctypes.data, m, int(bool(full_ortho)),
                                0.0 if eps is None else float(eps), alphas.ctypes.data, betas.ctypes.data,
                                E.ctypes.data, V.ctypes.data, None if Q is None else Q.ctypes.data,
                                ctypes.byref(used)))
    u = used.value
    return (alphas[:u], betas[:max(0, u - 1)], E[:u], V[:u * u].reshape(u, u, order="F"), Q, u, dt)


def lanczos_full(A, v0, m, full_ortho=False, out=None, eps=None):
    """lanczos_full(A, v0, m; full_ortho, out, eps) -> (E, V, Q_T) -- tools.jl:2742-2809."""
    _, _, E, V, Q, u, _ = _run(A, v0, m, full_ortho, eps, True, out)
    return E, V, (Q if u == Q.shape[0] else Q[:u])


class _LanczosVectors:
    """Lanczos rows regenerated from the stored coefficients (tools.jl:2852-2883)."""

    def __init__(self, op, initial, alphas, betas, dtype):
        self.op, self.initial, self.alphas, self.betas, self.dtype = op, initial, alphas, betas, dtype

    def __len__(self):
        return len(self.alphas)

    def __iter__(self):
        q = np.asarray(self.initial).astype(self.dtype)
        q = q / np.linalg.norm(q)
        yield q.copy()
        q_prev = np.zeros_like(q)
        for index in range(1, len(self.alphas)):
            res = self.op.matvec(q).astype(self.dtype)
            if index > 1:
                res = res - self.betas[index - 2] * q_prev
            res = res - self.alphas[index - 1] * q
            q_next = res / self.betas[index - 1]
            yield q_next.copy()
            q_prev, q = q, q_next


def lanczos_iter(A, v0, m, return_vec_iter=True, copy_v0=True, copy_A=False, eps=None):
    """lanczos_iter(A, v0, m; return_vec_iter, copy_v0, copy_A, eps) -- tools.jl:2885-2910."""
    alphas, betas, E, V, _, _, dt = _run(A, v0, m, False, eps, False)
    if not return_vec_iter:
        return E, V
    initial = np.array(v0, copy=True) if copy_v0 else v0
    return E, V, _LanczosVectors(A, initial, alphas, betas, dt)


def lin_comb_Q_T(coeff, Q_T, out=None):
    """lin_comb_Q_T(coeff, Q_T; out) -- tools.jl:2912-2944."""
    coeff = np.asarray(coeff)
    if isinstance(Q_T, np.ndarray):
        if coeff.shape[0] != Q_T.shape[0]:
            raise DimensionMismatch("one coefficient is required per Lanczos vector")
        res = Q_T.T @ coeff
        if out is not None:
            if out.shape != res.shape:
                raise DimensionMismatch("out must match a Lanczos vector")
            out[...] = res
            return out
        return res
    it = iter(Q_T)
    try:
        first = next(it)
    except StopIteration:
        raise ArgumentError("Lanczos vector iterator must be nonempty")
    if coeff.size == 0:
        raise DimensionMismatch("one coefficient is required per Lanczos vector")
    res = coeff[0] * first
    index = 0
    for vec in it:
        index += 1
        if index >= coeff.size:
            raise DimensionMismatch("too many Lanczos vectors")
        res = res + coeff[index] * vec
    if index + 1 != coeff.size:
        raise DimensionMismatch("too few Lanczos vectors")
    if out is not None:
        out[...] = res
        return out
    return res


def expm_lanczos(E, V, Q_T, a=1.0, out=None):
    """expm_lanczos(E, V, Q_T; a, out) -- tools.jl:2946-2952."""
    E = np.asarray(E)
    V = np.asarray(V)
    if not (E.shape[0] == V.shape[0] == V.shape[1]):
        raise DimensionMismatch("E and V dimensions must agree")
    coefficients = V @ (np.exp(a * E) * V[0, :])
    return lin_comb_Q_T(coefficients, Q_T, out=out)
