"""GPU parity of the batched engine: BlockOps / block_expm (tools.jl:2303-2718),
the lockstep batch (qs_batch_*), and sector <-> parent projection.

Bars: the reference's own BlockOps tolerances (runtests.jl:276-325: 5e-15
against the full evolve for a diagonal generator); Krylov results within
1e-10 of dense exponentials of the oracle's parent Hamiltonian; a batched
job bitwise identical alone and in a mix; projection round trips within
1e-13.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
import qsgpu

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _xxz(L, J=1.0, delta=2.0):
    b = [(i, i % L + 1) for i in range(1, L + 1)]
    return [("+-", [(J / 2, i, j) for i, j in b]), ("-+", [(J / 2, i, j) for i, j in b]),
            ("zz", [(J * delta, i, j) for i, j in b])]


def _colex_rank(state):
    r, k, p = 0, 1, 0
    while state:
        if state & 1:
            r += math.comb(p, k)
            k += 1
        state >>= 1
        p += 1
    return r


def test_blockops_particle_blocks_match_full_evolve(gpu):
    # runtests.jl:276-325 on the GPU engine: blocks over nup inside a parent without nup
    print("blockops nup blocks", flush=True)
    terms = [("z", [(1.0, 1), (0.5, 2)])]
    blocks = [dict(nup=n) for n in range(3)]
    ops = qsgpu.BlockOps(blocks, terms, [], qsgpu.SpinBasis1D, (2,), np.complex128, basis_kwargs=dict(pauli=False))
    assert len(ops.basis_dict) == 3
    assert len(ops.H_dict) == 0
    ops.compute_all_blocks()
    assert len(ops.H_dict) == 3
    full = oracle.Operator(oracle.Sector(2, nup=None, pauli=False), terms, check_symm=False).dense()
    psi = np.array([1, 2, 3, 4], dtype=np.complex128)
    psi /= np.linalg.norm(psi)
    times = [0.0, 0.2, 1.0]
    got = ops.evolve(psi, 0.0, times)
    want = np.stack([sl.expm(-1j * t * full) @ psi for t in times], axis=1)
    assert np.abs(got - want).max() <= 5e-15
    one = qsgpu.block_expm(ops, psi)
    assert np.abs(one - sl.expm(-1j * full) @ psi).max() <= 5e-15
    grid = qsgpu.block_expm(ops, psi, start=0.0, stop=1.0, num=3)
    assert np.abs(grid[:, 0] - psi).max() <= 3e-16
    assert np.abs(grid[:, -1] - sl.expm(-1j * full) @ psi).max() <= 5e-15


@pytest.mark.parametrize("L", [8, 10])
def test_blockops_momentum_blocks_against_dense(gpu, L):
    # the L=24 block-evolution benchmark in miniature: every momentum block of the half-filled XXZ chain (Delta = 2),
    # the Neel state projected into each, exp(-iHt) recombined in the parent space
    print("blockops momentum L=%d" % L, flush=True)
    nup = L // 2
    terms = _xxz(L)
    blocks = [dict(kblock=k) for k in range(L)]
    ops = qsgpu.BlockOps(blocks, terms, [], qsgpu.SpinBasis1D, (L,), np.complex128,
                         basis_kwargs=dict(nup=nup, pauli=False))
    par = oracle.Sector(L, nup=nup, pauli=False)
    H = oracle.Operator(par, terms, check_symm=False).dense()
    neel = sum(1 << i for i in range(0, L, 2))
    psi = np.zeros(math.comb(L, nup), dtype=np.complex128)
    psi[_colex_rank(neel)] = 1.0
    assert par.parent_states()[_colex_rank(neel)] == neel
    times = np.linspace(0.0, 4.0, 9)
    got = ops.evolve(psi, 0.0, times)
    w, V = np.linalg.eigh(H)
    want = np.stack([V @ (np.exp(-1j * w * t) * (V.conj().T @ psi)) for t in times], axis=1)
    assert np.abs(got - want).max() <= 1e-10
    # imaginary time renormalises nothing here (matches the reference's block evolve)
    got_im = ops.evolve(psi, 0.0, [0.3], imag_time=True)
    want_im = V @ (np.exp(-0.3 * w) * (V.conj().T @ psi))
    assert np.abs(got_im[:, 0] - want_im).max() <= 1e-10


def test_projection_round_trip(gpu):
    L, nup = 10, 5
    print("projection round trip", flush=True)
    rng = np.random.default_rng(3)
    n = math.comb(L, nup)
    psi = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), dtype=torch.complex128, device="cuda")
    acc = torch.zeros((1, n), dtype=torch.complex128, device="cuda")
    for k in range(L):
        b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, kblock=k)
        x = qsgpu.blocks.project(b, psi)
        # P' P = I on the block
        y = torch.zeros((1, n), dtype=torch.complex128, device="cuda")
        qsgpu.blocks.expand(b, x.reshape(1, -1).contiguous(), y, accumulate=False)
        x2 = qsgpu.blocks.project(b, y.reshape(-1))
        assert float((x2 - x).abs().max()) <= 1e-13
        qsgpu.blocks.expand(b, x.reshape(1, -1).contiguous(), acc, accumulate=True)
    # sum_k P_k P_k' = I on the parent space
    assert float((acc.reshape(-1) - psi).abs().max()) <= 1e-13
    # against the oracle's explicit projector columns for one block
    b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, kblock=3)
    sec = oracle.Sector(L, nup=nup, pauli=False, kblock=3)
    P = sec.projector_dense()
    x = qsgpu.blocks.project(b, psi).cpu().numpy()
    assert np.abs(x - P.conj().T @ psi.cpu().numpy()).max() <= 1e-13


def test_batch_job_alone_equals_in_mix(gpu):
    # a job's result does not depend on what else is in the batch (bitwise)
    print("batch alone vs mix", flush=True)
    L, nup = 12, 6
    terms = _xxz(L, delta=1.5)
    bases = [qsgpu.SpinBasis1D(L, nup=nup, pauli=False, kblock=k) for k in (0, 1, 3, 6)]
    ops = [qsgpu.QuantumLinearOperator(b, terms) for b in bases]
    rng = np.random.default_rng(9)
    vs = [torch.tensor(rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns), dtype=torch.complex128,
                       device="cuda") for b in bases]
    scales = [-1j * np.array([0.1, 0.7, 2.5]), -1j * np.array([1.3]), -1j * np.array([0.0, 4.0]),
              -np.array([0.2, 0.9])]
    mix = qsgpu.Batch(ops)
    jobs = [(i, vs[i], scales[i]) for i in range(len(ops))]
    outs = mix.expm_multiply(jobs)
    for i in range(len(ops)):
        alone = qsgpu.Batch([ops[i]]).expm_multiply([(0, vs[i], scales[i])])[0]
        assert torch.equal(alone, outs[i]), i
        # and against the single-operator driver (same arithmetic, own launches)
        single = qsgpu.expm_multiply(ops[i], vs[i].cpu().numpy(), scales[i])
        assert np.abs(outs[i].cpu().numpy().T - single).max() <= 1e-12 * max(1.0, np.abs(single).max())
    # batched matvec against the single launches
    ys = [torch.zeros_like(v) for v in vs]
    mix.matvec([(i, vs[i], ys[i]) for i in range(len(ops))])
    for i in range(len(ops)):
        want = ops[i].matvec(vs[i].cpu().numpy())
        assert np.abs(ys[i].cpu().numpy() - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("cdt", ["c128", "c64"])
def test_large_batch_job_table_equals_alone(gpu, cdt):
    # more jobs than one launch's kernel arguments hold: vector operations and
    # assembled products run through a device job table, one launch each;
    # every job still gives exactly its alone result (ComplexF64, and the
    # ComplexF32 stream with FP64 reductions)
    print("batch job table", cdt, flush=True)
    ndt, tdt, tol = ((np.complex128, torch.complex128, 1e-13) if cdt == "c128"
                     else (np.complex64, torch.complex64, 1e-6))
    L, nup = 12, 6
    terms = _xxz(L, delta=1.5)
    bases = [qsgpu.SpinBasis1D(L, nup=nup, pauli=False, kblock=k) for k in (0, 1, 3, 6)]
    ops = [qsgpu.QuantumLinearOperator(b, terms).assemble() for b in bases]
    rng = np.random.default_rng(21)
    jobs = []
    for i in range(40):
        o = i % len(ops)
        v = torch.tensor(rng.standard_normal(bases[o].Ns) + 1j * rng.standard_normal(bases[o].Ns),
                         dtype=tdt, device="cuda")
        jobs.append((o, v, -1j * np.array([0.3 + 0.05 * i, 1.7])))
    mix = qsgpu.Batch(ops, ndt)
    outs = mix.expm_multiply(jobs)
    ys = [torch.zeros_like(v) for _, v, _ in jobs]
    mix.matvec([(o, v, y) for (o, v, _), y in zip(jobs, ys)])
    for i, (o, v, sc) in enumerate(jobs):
        alone = qsgpu.Batch([ops[o]], ndt).expm_multiply([(0, v, sc)])[0]
        assert torch.equal(alone, outs[i]), i
        y1 = torch.zeros_like(v)
        qsgpu.Batch([ops[o]], ndt).matvec([(0, v, y1)])
        assert torch.equal(y1, ys[i]), i
        want = ops[o].matvec(v.cpu().numpy().astype(np.complex128))
        assert np.abs(ys[i].cpu().numpy() - want).max() <= tol * np.abs(want).max()


def test_batch_fp32_stream_form(gpu):
    # ComplexF32-stored Krylov bases with FP64 reductions against the FP64 batch
    print("batch fp32", flush=True)
    L, nup = 12, 6
    terms = _xxz(L)
    bases = [qsgpu.SpinBasis1D(L, nup=nup, pauli=False, kblock=k) for k in (0, 5)]
    ops = [qsgpu.QuantumLinearOperator(b, terms) for b in bases]
    rng = np.random.default_rng(1)
    v64 = [torch.tensor(rng.standard_normal(b.Ns) + 0j, dtype=torch.complex128, device="cuda") for b in bases]
    sc = -1j * np.array([0.5, 1.5])
    o64 = qsgpu.Batch(ops, np.complex128).expm_multiply([(i, v64[i], sc) for i in range(2)])
    o32 = qsgpu.Batch(ops, np.complex64).expm_multiply([(i, v64[i].to(torch.complex64), sc) for i in range(2)],
                                                        tol=1e-7)
    for a, b in zip(o64, o32):
        err = float((a - b.to(torch.complex128)).abs().max() / a.abs().max())
        assert err <= 1e-5, err


def test_batch_rejects_general_couplings(gpu):
    b = qsgpu.SpinBasis1D(8, nup=4, pauli=False)
    op = qsgpu.QuantumLinearOperator(b, [("+z-", [(0.5, 1, 2, 3)]), ("-z+", [(0.5, 1, 2, 3)])], check_symm=False)
    with pytest.raises(qsgpu.ArgumentError, match="hop"):
        qsgpu.Batch([op])


@pytest.mark.parametrize("dtype,tol,bound", [(np.complex128, 1e-12, 1e-11), (np.complex64, 1e-6, 2e-5)],
                         ids=["c128", "c64"])
def test_blockops_long_times_deep_recursion(gpu, dtype, tol, bound):
    # t up to 20 at L=14 drives the halving recursion past depth 3 (split midpoints
    # feeding further splits) in every block; exact per-block evolution as the bar
    print("blockops long times", np.dtype(dtype).name, flush=True)
    L, nup = 14, 7
    terms = _xxz(L)
    ops = qsgpu.BlockOps([dict(kblock=k) for k in range(L)], terms, [], qsgpu.SpinBasis1D, (L,), np.complex128,
                         basis_kwargs=dict(nup=nup, pauli=False))
    neel = sum(1 << i for i in range(0, L, 2))
    n = math.comb(L, nup)
    psi = np.zeros(n, dtype=np.complex128)
    psi[_colex_rank(neel)] = 1.0
    times = np.array([0.0, 5.0, 10.0, 20.0])
    got = ops.evolve(psi, 0.0, times, dtype=dtype, tol=tol)
    exact = torch.zeros((len(times), n), dtype=torch.complex128, device="cuda")
    psig = torch.tensor(psi, device="cuda")
    for k in range(L):
        b = qsgpu.SpinBasis1D(L, nup=nup, pauli=False, kblock=k)
        x = qsgpu.blocks.project(b, psig).cpu().numpy()
        if np.linalg.norm(x) < 1e-12:
            continue
        w, V = np.linalg.eigh(qsgpu.QuantumLinearOperator(b, terms).toarray())
        st = np.stack([V @ (np.exp(-1j * w * t) * (V.conj().T @ x)) for t in times])
        qsgpu.blocks.expand(b, torch.tensor(st, device="cuda").contiguous(), exact, accumulate=True)
    err = np.abs(got - exact.cpu().numpy().T).max(axis=0)
    assert err.max() <= bound, err
    assert np.abs(np.linalg.norm(got, axis=0) - 1).max() <= bound


def test_host_blocks_evolve_matches_blockops(gpu):
    # qs_blocks_evolve (host vectors, the Julia binding's entry point) == BlockOps.evolve, bitwise
    L, nup = 10, 5
    terms = _xxz(L)
    ops = qsgpu.BlockOps([dict(kblock=k) for k in range(L)], terms, [], qsgpu.SpinBasis1D, (L,), np.complex128,
                         basis_kwargs=dict(nup=nup, pauli=False), compute_all_blocks=True)
    rng = np.random.default_rng(8)
    psi = rng.standard_normal(math.comb(L, nup)) + 1j * rng.standard_normal(math.comb(L, nup))
    psi /= np.linalg.norm(psi)
    times = np.array([0.0, 0.7, 3.0])
    want = ops.evolve(psi, 0.0, times)
    keys = list(ops.basis_dict)
    got = qsgpu.blocks.blocks_evolve([ops.H_dict[k] for k in keys], [0] * len(keys), psi, -1j * times)
    assert np.array_equal(got, want)


def test_block_evolve_many_times_on_matrix_cores(gpu):
    # 24 time points per block from the first factorization: the batched
    # recombination of every block runs on the FP64 matrix cores
    L = 12   # blocks of about 77 rows: whole 16-row tiles for the matrix cores
    print("block evolve 24 times", flush=True)
    nup = L // 2
    terms = _xxz(L)
    blocks = [dict(kblock=k) for k in range(L)]
    ops = qsgpu.BlockOps(blocks, terms, [], qsgpu.SpinBasis1D, (L,), np.complex128,
                         basis_kwargs=dict(nup=nup, pauli=False))
    par = oracle.Sector(L, nup=nup, pauli=False)
    H = oracle.Operator(par, terms, check_symm=False).dense()
    neel = sum(1 << i for i in range(0, L, 2))
    psi = np.zeros(math.comb(L, nup), dtype=np.complex128)
    psi[_colex_rank(neel)] = 1.0
    times = np.linspace(0.0, 0.5, 24)
    got = ops.evolve(psi, 0.0, times)
    w, V = np.linalg.eigh(H)
    want = np.stack([V @ (np.exp(-1j * w * t) * (V.conj().T @ psi)) for t in times], axis=1)
    assert np.abs(got - want).max() <= 1e-10
