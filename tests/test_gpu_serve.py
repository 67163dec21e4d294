"""GPU tests of the serving loop (qs_server_*): mixed eigsh / evolve requests.

Bars: every request's result is bitwise the same alone (one worker, one
request at a time) and in the mix (four overlapping workers, all requests
queued at once); eigenvalues within 1e-10 of dense results of the oracle and
evolved states within four times the error of the reference's own Krylov path
(its numpy restatement, oracle/krylov.py) on the same request; worker errors
reach the caller with the reference's exception class.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from oracle import krylov
import qsgpu
from qsgpu import serve

pytestmark = pytest.mark.gpu


def _requests(n=14):
    return serve.poisson_workload(n, rate=200.0, seed=7, Ls=(10, 12))


def _run(reqs, builder, n_workers, sequential):
    out = []
    with serve.Server(n_workers=n_workers, hbm_budget=8e9) as srv:
        tickets = []
        for r in reqs:
            op = builder.operator(r)
            if r["kind"] == "eigsh":
                t = srv.submit_eigsh(op, k=min(r["k"], op.Ns), which="SA", tol=1e-12)
            else:
                t = srv.submit_evolve(op, builder.initial_state(r, op), 0.0, [r["t"]])
            if sequential:
                out.append(srv.wait(t))
            else:
                tickets.append(t)
        for t in tickets:
            out.append(srv.wait(t))
    return out


def test_alone_equals_in_mix_and_matches_dense(gpu):
    print("serve alone vs mix", flush=True)
    reqs = _requests()
    builder = serve.RequestBuilder()
    alone = _run(reqs, builder, 1, sequential=True)
    mix = _run(reqs, builder, 4, sequential=False)
    for r, a, m in zip(reqs, alone, mix):
        if r["kind"] == "eigsh":
            assert np.array_equal(a["vals"], m["vals"]), r
        else:
            assert np.array_equal(a["states"], m["states"]), r
        assert m["seconds"][2] >= m["seconds"][1] >= 0
    # accuracy against the oracle's dense sector Hamiltonians
    for r, m in zip(reqs, mix):
        sec = oracle.Sector(r["L"], nup=r["nup"], pauli=False, **r["blocks"])
        H = oracle.Operator(sec, serve.chain_terms(r["L"], r["J2"], r["delta"]), check_symm=False).dense()
        if r["kind"] == "eigsh":
            w = np.linalg.eigvalsh(H)
            assert np.allclose(m["vals"], w[:min(r["k"], len(w))], atol=1e-10, rtol=0), r
        else:
            op = builder.operator(r)
            v = builder.initial_state(r, op)
            want = sl.expm(-1j * r["t"] * H) @ v
            # the reference's Hermitian Krylov path is Lanczos without
            # reorthogonalisation (operators.jl:1686-1700): on a sector no larger
            # than krylov_dim it runs to the full dimension, loses orthogonality
            # and drifts from exp(-iHt)v (1.1e-8 on the L=10 k=8 request of this
            # stream).  The bar is four times the drift of the numpy
            # restatement of that path (oracle/krylov.py) on the same request.
            ref = krylov.expmv_times(H, v, [-1j * r["t"]], hermitian=True)[:, 0]
            bound = max(1e-10, 4 * np.abs(ref - want).max())
            assert np.abs(m["states"][:, 0] - want).max() <= bound, r


def test_budget_admits_one_job_at_a_time(gpu):
    # a budget smaller than one job's working set still runs every job (alone)
    print("serve small budget", flush=True)
    reqs = _requests(6)
    builder = serve.RequestBuilder()
    full = _run(reqs, builder, 4, sequential=False)
    out = []
    with serve.Server(n_workers=4, hbm_budget=1) as srv:
        ts = []
        for r in reqs:
            op = builder.operator(r)
            if r["kind"] == "eigsh":
                ts.append(srv.submit_eigsh(op, k=min(r["k"], op.Ns), which="SA", tol=1e-12))
            else:
                ts.append(srv.submit_evolve(op, builder.initial_state(r, op), 0.0, [r["t"]]))
        out = [srv.wait(t) for t in ts]
    for r, a, b in zip(reqs, full, out):
        key = "vals" if r["kind"] == "eigsh" else "states"
        assert np.array_equal(a[key], b[key])


def test_errors_reach_the_caller(gpu):
    b = qsgpu.SpinBasis1D(12, nup=6, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, serve.chain_terms(12, 0.2, 1.0))
    with serve.Server(n_workers=2, hbm_budget=1e9) as srv:
        with pytest.raises(qsgpu.ArgumentError):
            srv.submit_eigsh(op, k=b.Ns + 1)
        t = srv.submit_eigsh(op, k=4, maxiter=1, tol=1e-14, ncv=9)
        with pytest.raises(qsgpu.ErrorException):
            srv.wait(t)
        # the server keeps serving after a failed request
        t = srv.submit_eigsh(op, k=1)
        res = srv.wait(t)
        assert res["vals"].shape == (1,)


def test_concurrent_jobs_on_one_operator(gpu):
    # one operator shared by overlapping jobs of different element types: a
    # parent-table sector of more than 2^17 rows, whose products pre-scale x
    # into a buffer per stream (L=26 nup=13 k=0: 400,024 rows).  Every result
    # equals the same job run alone.
    print("serve shared operator", flush=True)
    L = 26
    b = qsgpu.SpinBasis1D(L, nup=13, pauli=False, kblock=0)
    assert b.Ns > 1 << 17
    op = qsgpu.QuantumLinearOperator(b, serve.chain_terms(L, 0.3, 0.8))
    rng = np.random.default_rng(31)
    v0 = rng.standard_normal(b.Ns)
    v1 = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    v1 /= np.linalg.norm(v1)
    jobs = [("eigsh", np.float64, 2), ("evolve", np.complex128, 0.7), ("eigsh", np.complex128, 1),
            ("evolve", np.complex64, 0.4), ("eigsh", np.float32, 1), ("evolve", np.complex128, 1.3)]

    def submit(srv, job):
        kind, dt, arg = job
        if kind == "eigsh":
            return srv.submit_eigsh(op, k=arg, which="SA", tol=1e-10, v0=v0, dtype=dt)
        return srv.submit_evolve(op, v1, 0.0, [arg], dtype=dt)

    def key(job):
        return "vals" if job[0] == "eigsh" else "states"

    with serve.Server(n_workers=1, hbm_budget=50e9) as srv:
        alone = [srv.wait(submit(srv, j)) for j in jobs]
    with serve.Server(n_workers=4, hbm_budget=50e9) as srv:
        tickets = [submit(srv, j) for j in jobs]
        mix = [srv.wait(t) for t in tickets]
    for j, a, m in zip(jobs, alone, mix):
        assert np.array_equal(a[key(j)], m[key(j)]), j
