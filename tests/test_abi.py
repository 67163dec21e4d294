"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and host-side validation maps to the reference's errors."""
import ctypes
import os
import re

import numpy as np
import pytest

import qsgpu
from qsgpu import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "qsgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(qs_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = qsgpu.lib()
    names = header_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_library_is_gfx950_only():
    # the shared object carries a gfx950 code object and nothing for other targets
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
    for other in (b"amdhsa--gfx90a", b"amdhsa--gfx942", b"nvptx"):
        assert other not in data


@pytest.mark.parametrize("kwargs,msg", [
    (dict(L=0), "L must be between 1 and 63"),
    (dict(L=64), "L must be between 1 and 63"),
    (dict(L=8, nup=9), "nup must lie between 0 and L"),
    (dict(L=8, nup=4, kblock=1, pblock=1), "pblock can be combined with kblock only at k=0 or pi"),
    (dict(L=8, nup=3, zblock=1), "symmetry transformation leaves the selected particle sector"),
    (dict(L=8, nup=4, pblock=2), "pblock must be +1 or -1"),
    (dict(L=8, nup=4, kblock=0, a=3), "a must be a positive divisor of L"),
])
def test_basis_argument_errors(kwargs, msg):
    # basis.jl:652-665, 686-687, 717-723, 744-745; symmetry_basis.jl:48-50
    with pytest.raises(qsgpu.ArgumentError, match=re.escape(msg)):
        qsgpu.SpinBasis1D(**kwargs)


def test_operator_term_validation():
    # operators.jl:48-58
    with pytest.raises(qsgpu.ArgumentError):
        qsgpu.OperatorTerm("", [(1.0,)])
    with pytest.raises(qsgpu.ArgumentError):
        qsgpu.OperatorTerm("zz", [(1.0, 1)])
    with pytest.raises(qsgpu.ArgumentError):
        qsgpu.OperatorTerm("z", [(1.0, 1.5)])
    t = qsgpu.OperatorTerm("+-", [[0.5, 1, 2]])
    assert t.couplings == [(0.5, 1, 2)]


def test_last_error_roundtrip():
    lib = qsgpu.lib()
    h = ctypes.c_void_p()
    st = lib.qs_basis_create(70, -1, 1, 1, *([_lib.QS_UNSET] * 6), 0, ctypes.byref(h))
    assert st == _lib.QS_EARG
    assert "L must be between" in _lib.last_error()
    buf = ctypes.create_string_buffer(8)
    n = lib.qs_last_error(buf, 8)
    assert n > 8 and len(buf.value) == 7
