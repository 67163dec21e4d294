"""GPU test of qs_op_matvec_noalloc, the product that allocates no device memory.

It is called through the C ABI on device buffers and compared with
qs_op_matvec (the allocating entry point) on the same inputs: FP64 within
1e-12 relative (bitwise where both run the same kernel), FP32 within 2e-6;
two calls in a row, the second on a smaller sector, so the shared scratch
array holds the first call's values when the second starts; sizes beyond the
scratch maximum are rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

import _dense
import qsgpu
from qsgpu import _lib
from qsgpu.errors import check

pytestmark = pytest.mark.gpu

_TMAP = None


def _tmap():
    import torch
    return {torch.float64: _lib.QS_F64, torch.complex128: _lib.QS_C128, torch.float32: _lib.QS_F32,
            torch.complex64: _lib.QS_C64}


def _noalloc(op, x, y, alpha=1.0, beta=0.0, transposed=False, conjugated=False):
    a, b = complex(alpha), complex(beta)
    ncols = 1 if x.dim() == 1 else x.shape[1]
    return _lib.lib().qs_op_matvec_noalloc(op._h, _tmap()[x.dtype], ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(y.data_ptr()), ncols, a.real, a.imag, b.real, b.imag,
                                           int(transposed), int(conjugated))


def _ref(op, x, y0, alpha, beta, transposed=False, conjugated=False):
    y = y0.clone()
    a, b = complex(alpha), complex(beta)
    check(_lib.lib().qs_op_matvec(op._h, _tmap()[x.dtype], ctypes.c_void_p(x.data_ptr()),
                                  ctypes.c_void_p(y.data_ptr()), 1, a.real, a.imag, b.real, b.imag,
                                  int(transposed), int(conjugated), 1))
    return y


def _case(L, blocks, J2, seed, dtype, alpha, beta):
    import torch
    b = qsgpu.SpinBasis1D(L, nup=L // 2, pauli=False, **blocks)
    op = qsgpu.QuantumLinearOperator(b, _dense.xxz_terms(L, J=1.0, delta=0.8, J2=J2))
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn(b.Ns, generator=g, dtype=torch.float64).to(dtype).cuda()
    y0 = torch.randn(b.Ns, generator=g, dtype=torch.float64).to(dtype).cuda()
    want = _ref(op, x, y0, alpha, beta)
    y = y0.clone()
    check(_noalloc(op, x, y, alpha, beta))
    torch.cuda.synchronize()
    return b, op, want, y


def test_noalloc_matches_allocating_product_twice_then_smaller(gpu):
    import torch
    print("noalloc: L=28 k0 p+ z+ then L=24 k0 p+", flush=True)
    # first call: the C2 sector (361,270 rows, parent table, pre-scaled x in
    # the scratch); the allocating product runs the same pre-scaled kernel
    b1, op1, want1, y1 = _case(28, dict(kblock=0, pblock=1, zblock=1), 0.4, 1, torch.float64, 0.75, -0.5)
    assert b1.Ns * 8 <= _lib.NOALLOC_SCRATCH_BYTES
    assert torch.equal(y1, want1)
    # second call, smaller sector: the scratch still holds the first call's x
    b2, op2, want2, y2 = _case(24, dict(kblock=0, pblock=1), 0.3, 2, torch.float64, 1.0, 0.0)
    assert b2.Ns < b1.Ns
    ref = torch.linalg.vector_norm(want2).item()
    assert torch.linalg.vector_norm(y2 - want2).item() <= 1e-12 * ref
    # and the first operator again, complex FP32 stream
    import torch as t
    x = t.randn(b1.Ns, dtype=t.complex128).to(t.complex64).cuda()
    y0 = t.zeros_like(x)
    want = _ref(op1, x, y0, 1.0, 0.0)
    y = y0.clone()
    check(_noalloc(op1, x, y))
    t.cuda.synchronize()
    assert (t.linalg.vector_norm((y - want).to(t.complex128)) / t.linalg.vector_norm(want.to(t.complex128))).item() \
        <= 2e-6


def test_noalloc_flags_and_implicit_sectors(gpu):
    import torch
    print("noalloc: flags, implicit sector", flush=True)
    L = 20
    b = qsgpu.SpinBasis1D(L, nup=10, pauli=False, zblock=1)
    terms = [("+-", [(0.5 + 0.1j, i, i % L + 1) for i in range(1, L + 1)]),
             ("-+", [(0.5 - 0.1j, i, i % L + 1) for i in range(1, L + 1)]),
             ("zz", [(0.7, i, i % L + 1) for i in range(1, L + 1)])]
    op = qsgpu.QuantumLinearOperator(b, terms, check_symm=False)
    x = torch.randn(b.Ns, dtype=torch.complex128).cuda()
    y0 = torch.randn(b.Ns, dtype=torch.complex128).cuda()
    for tr, cj in ((False, False), (True, False), (False, True), (True, True)):
        want = _ref(op, x, y0, 0.5 + 0.25j, 1.5, tr, cj)
        y = y0.clone()
        check(_noalloc(op, x, y, 0.5 + 0.25j, 1.5, tr, cj))
        torch.cuda.synchronize()
        ref = torch.linalg.vector_norm(want).item()
        assert torch.linalg.vector_norm(y - want).item() <= 1e-12 * ref, (tr, cj)


def test_noalloc_rejects_what_it_cannot_do_without_allocating(gpu):
    import torch
    print("noalloc: rejections", flush=True)
    # a parent-table sector larger than the scratch: 30 sites, k=0 (about 5.2e6 rows)
    b = qsgpu.SpinBasis1D(30, nup=15, pauli=False, kblock=0)
    assert b.Ns * 8 > _lib.NOALLOC_SCRATCH_BYTES
    op = qsgpu.QuantumLinearOperator(b, _dense.xxz_terms(30, J=1.0, delta=1.0))
    x = torch.zeros(b.Ns, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    assert _noalloc(op, x, y) == _lib.QS_EARG
    assert "scratch" in _lib.last_error()
    # an aliased input needs the copy the reference makes
    b = qsgpu.SpinBasis1D(12, nup=6, pauli=False, kblock=0)
    op = qsgpu.QuantumLinearOperator(b, _dense.xxz_terms(12, J=1.0, delta=1.0))
    x = torch.ones(b.Ns, dtype=torch.float64, device="cuda")
    assert _noalloc(op, x, x) == _lib.QS_EARG
    assert "alias" in _lib.last_error()
