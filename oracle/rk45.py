"""numpy restatement of QuSpin.jl's evolution under a time-dependent Hamiltonian (test infrastructure).

Only ``tests/`` import this module, as the checker of the GPU's qs_evolve_dynamic.
It follows the reference step for step:

* ``rk45_interval`` -- _rk45_interval, operators.jl:1964-2071: Dormand-Prince
  5(4) stages, error norm max |fifth - fourth| / (atol + rtol max(|state|, |fifth|)),
  accept when <= 1, step factor clamp(0.9 err^(-1/5), 0.2, 5) (5 when err = 0);
* ``evolve``        -- the dynamic branch of evolve, operators.jl:2092-2128: one
  interval per target time, imaginary-time states renormalised per column.

H(t) = H0 + sum_j f_j(t, *args_j) H_j with dense numpy matrices.
"""
import numpy as np


def _eps(x):
    x = abs(float(x))
    return np.nextafter(x, np.inf) - x


def _derivative(H0, dyn, v, t, imag_time):
    left = H0 @ v
    for M, f, args in dyn:
        c = f(t, *args)
        if c == 0:
            continue
        left = left + c * (M @ v)
    return (-1.0 if imag_time else -1j) * left


def rk45_interval(H0, dyn, initial, start, target, initial_step, max_step, imag_time, rtol=1e-9, atol=1e-11):
    """operators.jl:1964-2071; returns (state, next step)."""
    state = initial
    time = float(start)
    step = min(max_step, max(initial_step, _eps(target + 1.0)))
    d = lambda v, t: _derivative(H0, dyn, v, t, imag_time)
    while time < target:
        remaining = target - time
        if remaining <= 8 * _eps(max(abs(time), abs(target), 1.0)):
            time = float(target)
            break
        step = min(step, remaining)
        if not step > _eps(max(abs(time), 1.0)):
            raise RuntimeError("adaptive evolution step underflow")
        k1 = d(state, time)
        k2 = d(state + step * (1 / 5) * k1, time + step * (1 / 5))
        k3 = d(state + step * ((3 / 40) * k1 + (9 / 40) * k2), time + step * (3 / 10))
        k4 = d(state + step * ((44 / 45) * k1 - (56 / 15) * k2 + (32 / 9) * k3), time + step * (4 / 5))
        k5 = d(state + step * ((19372 / 6561) * k1 - (25360 / 2187) * k2 + (64448 / 6561) * k3 - (212 / 729) * k4),
               time + step * (8 / 9))
        k6 = d(state + step * ((9017 / 3168) * k1 - (355 / 33) * k2 + (46732 / 5247) * k3 + (49 / 176) * k4
                               - (5103 / 18656) * k5), time + step)
        fifth = state + step * ((35 / 384) * k1 + (500 / 1113) * k3 + (125 / 192) * k4 - (2187 / 6784) * k5
                                + (11 / 84) * k6)
        k7 = d(fifth, time + step)
        fourth = state + step * ((5179 / 57600) * k1 + (7571 / 16695) * k3 + (393 / 640) * k4
                                 - (92097 / 339200) * k5 + (187 / 2100) * k6 + (1 / 40) * k7)
        scale = atol + rtol * np.maximum(np.abs(state), np.abs(fifth))
        error = float(np.max(np.abs(fifth - fourth) / scale, initial=0.0))
        if error <= 1:
            state = fifth
            time += step
        factor = 5.0 if error == 0 else min(5.0, max(0.2, 0.9 * error ** (-1 / 5)))
        step = min(max_step, step * factor)
    return state, step


def evolve(H0, dyn, v0, t0, times, imag_time=False, max_step=0.01, rtol=1e-9, atol=1e-11):
    """operators.jl:2092-2128: states at each target time (columns / last axis)."""
    state = np.asarray(v0, dtype=np.complex128).copy()
    current = float(t0)
    next_step = max_step
    out = []
    for target in times:
        if not target >= current:
            raise ValueError("times must be sorted and not precede t0")
        state, next_step = rk45_interval(H0, dyn, state, current, float(target), next_step, max_step, imag_time,
                                         rtol, atol)
        current = float(target)
        if imag_time:
            state = state / np.sqrt(np.sum(np.abs(state) ** 2, axis=0))
        out.append(state.copy())
    return np.stack(out, axis=-1)
