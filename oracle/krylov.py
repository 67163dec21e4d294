"""numpy restatement of QuSpin.jl's Krylov exponential action (test infrastructure).

Only ``tests/`` import this module, as the checker of the GPU Krylov solvers.
It follows the reference step for step, on a callable ``matvec`` (or a dense
matrix), in FP64 / ComplexF64:

* ``factorization``  -- _krylov_factorization, operators.jl:1655-1731: three-term
  Lanczos without reorthogonalisation for a Hermitian operator, Arnoldi with
  two passes of Gram-Schmidt otherwise; breakdown when the new residual norm
  is <= 100 eps max(1, ||H_so_far||_F).
* ``project``        -- _krylov_project, operators.jl:1733-1756: beta V exp(s H_c) e_1
  and the a-posteriori estimate |s| h_{m+1,m} beta |exp(s H_c)[m, 1]|.
* ``expmv_vector``   -- _krylov_expmv_vector, operators.jl:1758-1803: the halving
  recursion (depth < 14, tolerance halved per level).
* ``expmv_times``    -- _krylov_expmv_times, operators.jl:1805-1893: one factorisation
  for every scale; a Hermitian factorisation diagonalised once.

Julia's ``exp(::Matrix)`` and ``eigen(::Hermitian)`` are replaced by
scipy.linalg.expm and numpy.linalg.eigh, which agree with them to rounding.
Pins: tests/test_oracle.py checks these functions against dense exponentials
and against the reference's own known answers.
"""
import numpy as np
import scipy.linalg as sl

_EPS = np.finfo(np.float64).eps


def _as_matvec(op):
    if callable(op):
        return op
    A = op
    return lambda v: A @ v


def factorization(op, v, hermitian, krylov_dim=30, dtype=np.complex128):
    """operators.jl:1655-1731."""
    mv = _as_matvec(op)
    initial = np.asarray(v, dtype=dtype).copy()
    beta = np.linalg.norm(initial)
    n = initial.size
    steps = min(int(krylov_dim), n)
    V = np.zeros((n, steps + 1), dtype=dtype)
    H = np.zeros((steps + 1, steps), dtype=dtype)
    V[:, 0] = initial / beta
    completed = steps
    terminal = 0.0
    if hermitian:
        previous = 0.0
        for col in range(steps):
            r = np.asarray(mv(V[:, col]), dtype=dtype).copy()
            if col > 0:
                r -= previous * V[:, col - 1]
            c = np.vdot(V[:, col], r)
            H[col, col] = c
            r -= c * V[:, col]
            terminal = np.linalg.norm(r)
            H[col + 1, col] = terminal
            if col < steps - 1:
                H[col, col + 1] = terminal
            if terminal <= 100 * _EPS * max(1.0, np.linalg.norm(H)):
                completed = col + 1
                break
            elif col < steps - 1:
                V[:, col + 1] = r / terminal
            previous = terminal
    else:
        for col in range(steps):
            r = np.asarray(mv(V[:, col]), dtype=dtype).copy()
            for row in range(col + 1):
                c = np.vdot(V[:, row], r)
                H[row, col] = c
                r -= c * V[:, row]
            for row in range(col + 1):
                c = np.vdot(V[:, row], r)
                H[row, col] += c
                r -= c * V[:, row]
            terminal = np.linalg.norm(r)
            H[col + 1, col] = terminal
            if terminal <= 100 * _EPS * max(1.0, np.linalg.norm(H)):
                completed = col + 1
                break
            elif col < steps - 1:
                V[:, col + 1] = r / terminal
    return dict(initial=initial, beta=beta, V=V, H=H, completed=completed, terminal=terminal, dimension=n,
                steps=steps, hermitian=hermitian)


def project(f, scale):
    """operators.jl:1733-1756: (state, estimate)."""
    c = f["completed"]
    E = sl.expm(scale * f["H"][:c, :c])
    state = f["beta"] * (f["V"][:, :c] @ E[:, 0])
    est = 0.0
    if c == f["steps"] and c < f["dimension"]:
        est = abs(scale) * f["terminal"] * f["beta"] * abs(E[-1, 0])
    return state, est


def expmv_vector(op, v, scale, hermitian, tol=1e-12, krylov_dim=30, depth=0):
    """operators.jl:1758-1803."""
    initial = np.asarray(v, dtype=np.complex128).copy()
    if scale == 0 or np.linalg.norm(initial) == 0:
        return initial
    f = factorization(op, initial, hermitian, krylov_dim)
    state, est = project(f, scale)
    if est > tol * max(1.0, np.linalg.norm(state)):
        if depth >= 14:
            raise RuntimeError("Krylov exponential action did not converge; increase krylov_dim")
        mid = expmv_vector(op, initial, scale / 2, hermitian, tol / 2, krylov_dim, depth + 1)
        return expmv_vector(op, mid, scale / 2, hermitian, tol / 2, krylov_dim, depth + 1)
    return state


def expmv_times(op, v, scales, hermitian, tol=1e-12, krylov_dim=30):
    """operators.jl:1805-1893: columns exp(s_j op) v."""
    initial = np.asarray(v, dtype=np.complex128).copy()
    out = np.empty((initial.size, len(scales)), dtype=np.complex128)
    if len(scales) == 0:
        return out
    if np.linalg.norm(initial) == 0 or all(s == 0 for s in scales):
        out[:] = initial[:, None]
        return out
    f = factorization(op, initial, hermitian, krylov_dim)
    c = f["completed"]
    if hermitian:
        w, U = np.linalg.eigh(f["H"][:c, :c])
        first = U.conj().T[:, 0]
        for j, s in enumerate(scales):
            if s == 0:
                out[:, j] = initial
                continue
            coef = U @ (np.exp(s * w) * first)
            state = f["beta"] * (f["V"][:, :c] @ coef)
            est = 0.0
            if c == f["steps"] and c < f["dimension"]:
                est = abs(s) * f["terminal"] * f["beta"] * abs(coef[-1])
            if est > tol * max(1.0, np.linalg.norm(state)):
                state = expmv_vector(op, initial, s, hermitian, tol, krylov_dim)
            out[:, j] = state
    else:
        for j, s in enumerate(scales):
            if s == 0:
                out[:, j] = initial
                continue
            state, est = project(f, s)
            if est > tol * max(1.0, np.linalg.norm(state)):
                state = expmv_vector(op, initial, s, hermitian, tol, krylov_dim)
            out[:, j] = state
    return out
