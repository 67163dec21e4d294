"""GPU tests of the weight-table form of assembled rows.

An assembled symmetry sector holds few distinct weights, so each stored entry
keeps a 16-bit index into a table of (re, im) pairs; QS_SELL_TABLE=0 at
assembly keeps the per-entry weights.  The two forms hold the same weights
in the same order, so every product (plain, the fused Lanczos form through
expm_multiply, and the batched one) and the exported CSR agree bit for bit,
and both match the oracle's sector operator.
"""
import os

import numpy as np
import pytest

import oracle
import qsgpu
from qsgpu import serve

pytestmark = pytest.mark.gpu


def _assembled(b, terms, table):
    old = os.environ.get("QS_SELL_TABLE")
    os.environ["QS_SELL_TABLE"] = "1" if table else "0"
    try:
        op = qsgpu.QuantumLinearOperator(b, terms).assemble()
    finally:
        if old is None:
            del os.environ["QS_SELL_TABLE"]
        else:
            os.environ["QS_SELL_TABLE"] = old
    op.set_variant(qsgpu.VARIANT_SELL)
    return op


@pytest.mark.parametrize("L,blocks", [(18, dict(kblock=3)), (18, dict(kblock=0, pblock=1, zblock=1)),
                                      (16, dict()), (20, dict(kblock=5, zblock=-1))])
def test_table_rows_match_per_entry_rows(gpu, L, blocks):
    print("sell table", L, blocks, flush=True)
    terms = serve.chain_terms(L, 0.3, 0.7)
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False, **blocks)
    tab = _assembled(b, terms, True)
    per = _assembled(b, terms, False)
    rng = np.random.default_rng(L)
    x = rng.standard_normal(b.Ns) + 1j * rng.standard_normal(b.Ns)
    for dt in (np.complex128, np.complex64):
        a, c = tab.matvec(x.astype(dt)), per.matvec(x.astype(dt))
        assert a.dtype == c.dtype and np.array_equal(a.view(np.uint8), c.view(np.uint8))
    ca, cb = tab.tocsr(), per.tocsr()
    assert np.array_equal(ca.indptr, cb.indptr) and np.array_equal(ca.indices, cb.indices)
    assert np.array_equal(ca.data.view(np.uint8), cb.data.view(np.uint8))
    v = x / np.linalg.norm(x)
    ya = qsgpu.expm_multiply(tab, v, np.array([-3.0j]))
    yb = qsgpu.expm_multiply(per, v, np.array([-3.0j]))
    assert np.array_equal(ya.view(np.uint8), yb.view(np.uint8))
    if L <= 18:
        H = oracle.Operator(oracle.Sector(L, nup=L // 2, pauli=False, **blocks), terms)
        want = H.matvec(x)
        assert np.linalg.norm(tab.matvec(x) - want) <= 1e-13 * np.linalg.norm(want)


def test_table_rows_in_the_batch_engine(gpu):
    # block evolution through the lockstep engine: every block's rows in the
    # table form against the per-entry form (blocks above the single-workgroup
    # size, so the batched products read the rows)
    import torch
    print("sell table batch", flush=True)
    L = 20
    terms = serve.chain_terms(L, 0.25, 1.2)
    bases = [qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False, kblock=k) for k in (1, 2, 3)]
    assert min(bb.Ns for bb in bases) > 2048
    outs = []
    for table in (True, False):
        ops = [_assembled(bb, terms, table) for bb in bases]
        batch = qsgpu.blocks.Batch(ops, np.complex128)
        jobs = []
        for i, bb in enumerate(bases):
            v = np.random.default_rng(i).standard_normal(bb.Ns) + 0j
            v = torch.from_numpy(v / np.linalg.norm(v)).to("cuda")
            jobs.append((i, v, -1j * np.linspace(0.5, 4.0, 5)))
        outs.append([s.cpu().numpy() for s in batch.expm_multiply(jobs, tol=1e-12)])
    for a, c in zip(*outs):
        assert np.array_equal(a.view(np.uint8), c.view(np.uint8))


def test_table_overflow_keeps_per_entry_weights(gpu):
    # all-pairs zz couplings with random strengths give nearly every state its
    # own diagonal value (far more than the table's 65,536 slots): the rows
    # keep their per-entry weights and the products stay exact
    print("sell table overflow", flush=True)
    L = 20
    rng = np.random.default_rng(7)
    pairs = [(float(rng.uniform(-1, 1)), i, j) for i in range(1, L + 1) for j in range(i + 1, L + 1)]
    terms = serve.chain_terms(L, 0.0, 0.0) + [("zz", pairs)]
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False)
    assert b.Ns > 65536
    tab = _assembled(b, terms, True)
    per = _assembled(b, terms, False)
    free = qsgpu.QuantumLinearOperator(b, terms)
    x = rng.standard_normal(b.Ns)
    a, c, f = tab.matvec(x), per.matvec(x), free.matvec(x)
    assert np.array_equal(a.view(np.uint8), c.view(np.uint8))
    assert np.linalg.norm(a - f) <= 1e-13 * np.linalg.norm(f)
