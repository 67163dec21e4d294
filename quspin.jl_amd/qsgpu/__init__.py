"""qsgpu -- MI355X host interface for QuSpin.jl's spin-1/2 sector Hamiltonian action.

This package mirrors the part of QuSpin.jl's public API that sits on the
matrix-free path (``SpinBasis1D``, ``SpinBasisGeneral``, ``OperatorTerm``, ``QuantumLinearOperator``,
``eigsh``, ``evolve``, ``ExpmMultiplyParallel``) with the same argument names,
meanings and exception classes, and forwards every call through the C ABI of
``libqsgpu.so`` (include/qsgpu.h) to hand-written gfx950 kernels.  There is no
CPU fallback: without the built library, importing the operator classes
works but every call raises ImportError.
"""
from ._lib import build, lib, LIB_PATH, VARIANT_HOP, VARIANT_SPLIT, VARIANT_SELL  # noqa: F401
from .errors import ArgumentError, DimensionMismatch, ErrorException  # noqa: F401
from .basis import SpinBasis1D, SpinBasisGeneral  # noqa: F401
from .operators import (  # noqa: F401
    OperatorTerm,
    QuantumLinearOperator,
    Hamiltonian,
    eigsh,
    evolve,
    expm_multiply,
    ExpmMultiplyParallel,
    apply,
    mul,
    device_info,
)
from .blocks import BlockOps, Batch, block_expm  # noqa: F401
from .serve import Server  # noqa: F401
from .lanczos import lanczos_full, lanczos_iter, lin_comb_Q_T, expm_lanczos  # noqa: F401
