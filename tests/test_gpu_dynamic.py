"""GPU tests of evolve with dynamic terms (qs_evolve_dynamic).

H(t) = H0 + sum_j f_j(t) H_j with every product on the device, against the
numpy restatement of the reference's RK45 (oracle/rk45.py, operators.jl:
1964-2071, 2092-2128) on the oracle's dense sector matrices (within 1e-10:
the same step sequence, rounding apart) and against scipy's DOP853 at tight
tolerances (within 2e-8, the scale of the reference's rtol = 1e-9).
"""
import numpy as np
import pytest

import _dense
import oracle
from oracle import rk45
import qsgpu

pytestmark = pytest.mark.gpu


def _setup(L, nup, blocks):
    static = _dense.xxz_terms(L, J=1.0, delta=0.6)
    stag = stag_of(L)
    hop = ("+-", [(0.3, i, i % L + 1) for i in range(1, L + 1)])
    hop2 = ("-+", [(0.3, i, i % L + 1) for i in range(1, L + 1)])
    f1 = lambda t, w: np.cos(w * t)
    f2 = lambda t: 0.5 * np.sin(t) ** 2
    sec = oracle.Sector(L, nup=nup, pauli=False, **blocks)
    H0 = oracle.Operator(sec, static, check_symm=False).dense()
    D1 = oracle.Operator(sec, [stag], check_symm=False).dense()
    D2 = oracle.Operator(sec, [hop, hop2], check_symm=False).dense()
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, **blocks)
    H = qsgpu.Hamiltonian(static, [(stag[0], stag[1], f1, (2.0,)), (hop[0], hop[1], f2, ()),
                                   (hop2[0], hop2[1], f2, ())], basis=b, check_symm=False)
    dyn = [(D1, f1, (2.0,)), (D2, f2, ())]
    return b, H, H0, dyn


@pytest.mark.parametrize("case", [(10, 5, {}), (12, 6, dict(kblock=0, pblock=1)), (12, 6, dict(zblock=1))],
                         ids=["L10", "L12_k0_p1", "L12_z1"])
def test_dynamic_evolve_matches_the_reference_rk45(gpu, case):
    from scipy.integrate import solve_ivp
    L, nup, blocks = case
    print("dynamic evolve", case, flush=True)
    b, H, H0, dyn = _setup(L, nup, blocks)
    rng = np.random.default_rng(31)
    v = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    v /= np.linalg.norm(v)
    times = [0.3, 1.1, 2.0]
    got = qsgpu.evolve(H, v, 0.0, times, max_step=0.05)
    want = rk45.evolve(H0, dyn, v, 0.0, times, max_step=0.05)
    assert np.abs(got - want).max() <= 1e-10
    assert qsgpu.evolve.last_n_steps > 0
    rhs = lambda t, y: -1j * (H0 @ y + sum(f(t, *a) * (M @ y) for M, f, a in dyn))
    sol = solve_ivp(rhs, (0.0, 2.0), v, method="DOP853", t_eval=times, rtol=1e-13, atol=1e-13)
    assert np.abs(got - sol.y).max() <= 2e-8
    # the ComplexF32-stream form follows to FP32 accuracy
    got32 = qsgpu.evolve(H, v, 0.0, times, max_step=0.05, dtype=np.complex64)
    assert np.abs(got32 - want).max() <= 1e-4


def test_dynamic_evolve_imaginary_time_matrix_input_and_errors(gpu):
    print("dynamic evolve imag / matrix", flush=True)
    b, H, H0, dyn = _setup(10, 5, {})
    rng = np.random.default_rng(32)
    V = rng.standard_normal((b.Ns, 2)) + 1j * rng.standard_normal((b.Ns, 2))
    got = qsgpu.evolve(H, V, 0.0, [0.5, 1.0], imag_time=True, max_step=0.05)
    want = rk45.evolve(H0, dyn, V, 0.0, [0.5, 1.0], imag_time=True, max_step=0.05)
    assert got.shape == (b.Ns, 2, 2)
    assert np.abs(got - want).max() <= 1e-10
    assert np.allclose(np.linalg.norm(got, axis=0), 1.0, atol=1e-12)
    # scalar time: one state
    s = qsgpu.evolve(H, V[:, 0], 0.0, 0.5, max_step=0.05)
    assert s.shape == (b.Ns,)
    with pytest.raises(qsgpu.ArgumentError, match="sorted"):
        qsgpu.evolve(H, V[:, 0], 0.0, [1.0, 0.5])
    with pytest.raises(qsgpu.ArgumentError, match="max_step"):
        qsgpu.evolve(H, V[:, 0], 0.0, [1.0], max_step=0.0)
    # without dynamic terms the Hamiltonian takes the Krylov exponential
    Hs = qsgpu.Hamiltonian(_dense.xxz_terms(10, J=1.0, delta=0.6), [], basis=b)
    import scipy.linalg as sl
    out = qsgpu.evolve(Hs, V[:, 0], 0.0, [0.7])
    assert np.abs(out[:, 0] - sl.expm(-0.7j * H0) @ V[:, 0]).max() <= 3e-11
    # static_fmt=:csr assembles the static part; products read the stored rows
    Hc = qsgpu.Hamiltonian(_dense.xxz_terms(10, J=1.0, delta=0.6), [(stag_of(10)[0], stag_of(10)[1], np.cos, ())],
                           basis=b, static_fmt="csr", check_symm=False)
    assert np.abs(Hc.apply(V[:, 0], 0.4) - (H0 + np.cos(0.4) * dyn[0][0]) @ V[:, 0]).max() <= 1e-12
    with pytest.raises(qsgpu.ArgumentError, match="static_fmt"):
        qsgpu.Hamiltonian(_dense.xxz_terms(10, J=1.0, delta=0.6), [], basis=b, static_fmt="coo")


def stag_of(L):
    return ("z", [((-1.0) ** i * 0.5, i) for i in range(1, L + 1)])
