"""BlockOps: many symmetry blocks of one parent space, evolved together.

Mirrors tools.jl:2303-2718 (BlockOps, compute_all_blocks!, update_blocks!,
evolve(::BlockOps), block_expm).  The reference evolves block after block
(or on `n_jobs` host threads); here the blocks' Krylov recursions advance in
lockstep through the batched engine of libqsgpu (qs_batch_*): one launch per
Krylov step for all blocks, every block's result bitwise independent of the
others.  Projection P' psi_0 and recombination sum_b P_b states_b run on the
device (qs_basis_project / qs_basis_expand).
"""
import ctypes
import numbers

import numpy as np

from . import _lib
from .basis import SpinBasis1D
from .errors import ArgumentError, DimensionMismatch, check
from .operators import QuantumLinearOperator

_EPS = np.finfo(np.float64).eps


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("BlockOps needs a GPU")
    return torch


_TDT = None


def _tdtype(dtype):
    torch = _torch()
    m = {np.dtype(np.complex128): (torch.complex128, _lib.QS_C128),
         np.dtype(np.complex64): (torch.complex64, _lib.QS_C64)}
    dt = np.dtype(dtype)
    if dt not in m:
        raise ArgumentError("block evolution needs a complex dtype (complex128 or complex64)")
    return m[dt]


class Batch:
    """A set of operators advanced in lockstep (qs_batch_create); device tensors in, device tensors out."""

    def __init__(self, ops, dtype=np.complex128):
        self.ops = list(ops)
        self.torch_dtype, self.code = _tdtype(dtype)
        arr = (ctypes.c_void_p * len(self.ops))(*[op._h for op in self.ops])
        h = ctypes.c_void_p()
        check(_lib.lib().qs_batch_create(arr, len(self.ops), self.code, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().qs_batch_destroy(h)
            except (TypeError, AttributeError):
                pass   # interpreter shutdown: module globals are already gone
            self._h = None

    def matvec(self, jobs):
        """jobs: list of (op_index, x, y) device tensors; y_j = op x_j."""
        torch = _torch()
        torch.cuda.synchronize()
        n = len(jobs)
        idx = np.asarray([j[0] for j in jobs], dtype=np.int32)
        xs = (ctypes.c_void_p * n)(*[j[1].data_ptr() for j in jobs])
        ys = (ctypes.c_void_p * n)(*[j[2].data_ptr() for j in jobs])
        check(_lib.lib().qs_batch_matvec(self._h, n, idx.ctypes.data_as(ctypes.c_void_p), xs, ys))

    def expm_multiply(self, jobs, tol=1e-12, krylov_dim=30):
        """jobs: list of (op_index, v, scales); returns per job a (len(scales), Ns) device tensor
        whose row j is exp(scales[j] * op) v (column-major Ns x n_times)."""
        torch = _torch()
        outs, scales, nts, idx, vs = [], [], [], [], []
        for op_i, v, sc in jobs:
            sc = np.atleast_1d(np.asarray(sc, dtype=np.complex128))
            ns = self.ops[op_i].Ns
            if v.shape != (ns,) or v.dtype != self.torch_dtype:
                raise DimensionMismatch("vector and operator dimensions (or dtype) do not match")
            outs.append(torch.zeros((sc.size, ns), dtype=self.torch_dtype, device=v.device))
            scales.append(sc)
            nts.append(sc.size)
            idx.append(op_i)
            vs.append(v.contiguous())
        torch.cuda.synchronize()
        n = len(jobs)
        sc_all = np.ascontiguousarray(np.concatenate(scales) if scales else np.zeros(1, np.complex128))
        nmv = np.zeros(max(1, n), dtype=np.int32)
        check(_lib.lib().qs_batch_expm_multiply(
            self._h, n, np.asarray(idx, dtype=np.int32).ctypes.data_as(ctypes.c_void_p),
            (ctypes.c_void_p * max(1, n))(*[v.data_ptr() for v in vs]),
            np.asarray(nts, dtype=np.int64).ctypes.data_as(ctypes.c_void_p), sc_all.ctypes.data_as(ctypes.c_void_p),
            float(tol), int(krylov_dim), (ctypes.c_void_p * max(1, n))(*[o.data_ptr() for o in outs]),
            nmv.ctypes.data_as(ctypes.c_void_p)))
        self.last_n_matvec = nmv[:n].tolist()
        return outs


_SYMMETRY_KWARGS = ("kblock", "pblock", "zblock", "pzblock", "zAblock", "zBblock")


def parent_size(basis, kind=0):
    n = ctypes.c_int64()
    check(_lib.lib().qs_basis_parent_size(basis._h, int(kind), ctypes.byref(n)))
    return n.value


def project(basis, parent, kind=0, dtype=np.complex128):
    """projector' * parent (tools.jl:2453) on device tensors.  kind 0: parent indexed like the
    basis' own particle sector (combinadic rank); kind 1: indexed by the state (2^L entries)."""
    torch = _torch()
    tdt, code = _tdtype(dtype)
    parent = parent.to(tdt).contiguous()
    if parent.shape != (parent_size(basis, kind),):
        raise DimensionMismatch("state and parent basis dimensions do not match")
    out = torch.empty(basis.Ns, dtype=tdt, device=parent.device)
    torch.cuda.synchronize()
    check(_lib.lib().qs_basis_project(basis._h, code, int(kind), ctypes.c_void_p(parent.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr())))
    return out


def expand(basis, x, parent, accumulate=True, kind=0, dtype=np.complex128):
    """parent (+)= projector * x -- tools.jl:2488-2495.  x: (ncols, Ns) and parent: (ncols, Npar)
    row-major tensors, i.e. the column-major Ns x ncols / Npar x ncols blocks the C ABI takes."""
    torch = _torch()
    tdt, code = _tdtype(dtype)
    ncols = 1 if x.dim() == 1 else x.shape[0]
    if x.dtype != tdt or parent.dtype != tdt or not x.is_contiguous() or not parent.is_contiguous():
        raise ArgumentError("x and parent must be contiguous tensors of the block dtype")
    if x.numel() != ncols * basis.Ns or parent.numel() != ncols * parent_size(basis, kind):
        raise DimensionMismatch("block, state and parent dimensions do not match")
    torch.cuda.synchronize()
    check(_lib.lib().qs_basis_expand(basis._h, code, int(kind), ctypes.c_void_p(x.data_ptr()),
                                     ctypes.c_void_p(parent.data_ptr()), ncols, int(bool(accumulate))))
    return parent


class BlockOps:
    """BlockOps(blocks, static, dynamic, basis_con, basis_args, dtype; ...) -- tools.jl:2303-2353."""

    def __init__(self, blocks, static, dynamic, basis_con, basis_args, dtype, basis_kwargs=None,
                 get_proj_kwargs=None, save_previous_data=True, compute_all_blocks=False, check_symm=True,
                 check_herm=True, check_pcon=True):
        if dynamic:
            raise ArgumentError("dynamic block Hamiltonians are not supported by the GPU block engine")
        self.basis_kwargs = dict(basis_kwargs or {})
        self.dtype = np.dtype(dtype)
        self.static = list(static)
        self.dynamic = []
        self.save_previous_data = bool(save_previous_data or compute_all_blocks)
        self.check_symm, self.check_herm, self.check_pcon = check_symm, check_herm, check_pcon
        self.basis_dict, self.H_dict = {}, {}
        self._basis_con, self._basis_args = basis_con, tuple(basis_args)
        if any(self.basis_kwargs.get(k) is not None for k in _SYMMETRY_KWARGS):
            raise ArgumentError("the parent basis of the GPU block engine carries no symmetry blocks")
        self._add_blocks(blocks, basis_con, basis_args)
        self._parent = basis_con(*self._basis_args, **self.basis_kwargs)
        if compute_all_blocks:
            self.compute_all_blocks()

    def _add_blocks(self, blocks, basis_con, basis_args):
        for block in blocks:
            key = repr(dict(block))
            if key in self.basis_dict:
                continue
            try:
                basis = basis_con(*basis_args, **{**self.basis_kwargs, **dict(block)})
            except ArgumentError as e:
                if "empty" in str(e):
                    continue
                raise
            if basis.Ns > 0:
                self.basis_dict[key] = basis

    def compute_all_blocks(self):
        """compute_all_blocks! -- tools.jl:2355-2378."""
        for key, basis in self.basis_dict.items():
            if key not in self.H_dict:
                self.H_dict[key] = self._hamiltonian(basis)
        return self

    def update_blocks(self, blocks, basis_con, basis_args, compute_all_blocks=False):
        """update_blocks! -- tools.jl:2380-2401."""
        self._add_blocks(blocks, basis_con, basis_args)
        if compute_all_blocks:
            self.compute_all_blocks()
        return self

    def _hamiltonian(self, basis):
        # the reference assembles every block (Hamiltonian(...; static_fmt=:csc),
        # tools.jl:2147-2168); the device assembly serves the lockstep engine's
        # products, and couplings it cannot assemble stay matrix-free
        op = QuantumLinearOperator(basis, self.static, check_symm=self.check_symm, check_herm=self.check_herm,
                                   check_pcon=self.check_pcon)
        try:
            op.assemble()
        except ArgumentError:
            pass
        return op

    def _block_data(self, key, basis):
        if key in self.H_dict:
            return self.H_dict[key]
        H = self._hamiltonian(basis)
        if self.save_previous_data:
            self.H_dict[key] = H
        return H

    def _kind(self, basis):
        # blocks over the particle number inside a parent without one: state-indexed parent
        return 1 if (basis.nup is not None and self._parent.nup is None) else 0

    def _entries(self, psi_0, dtype):
        """_active_block_entries -- tools.jl:2445-2475: blocks with |P' psi_0| > 1000 eps."""
        torch = _torch()
        tdt, _ = _tdtype(dtype)
        psi = torch.as_tensor(np.asarray(psi_0)).to(device="cuda", dtype=tdt)
        if psi.shape != (self._parent.Ns,):
            raise DimensionMismatch("initial state and parent basis dimensions do not match")
        entries = []
        for key, basis in self.basis_dict.items():
            H = self._block_data(key, basis)
            x = project(basis, psi, self._kind(basis), dtype)
            if float(torch.linalg.vector_norm(x.to(torch.complex128))) > 1000 * _EPS:
                entries.append((basis, H, x))
        if not entries:
            raise ArgumentError("initial state has no projection onto the selected blocks")
        return entries

    def _evolution(self, entries, scales, npar, dtype, tol, krylov_dim):
        torch = _torch()
        tdt, _ = _tdtype(dtype)
        batch = Batch([e[1] for e in entries], dtype)
        states = batch.expm_multiply([(i, e[2], scales) for i, e in enumerate(entries)], tol=tol,
                                     krylov_dim=krylov_dim)
        self.last_n_matvec = batch.last_n_matvec
        result = torch.zeros((len(scales), npar), dtype=tdt, device="cuda")
        for (basis, _, _), st in zip(entries, states):
            expand(basis, st, result, accumulate=True, kind=self._kind(basis), dtype=dtype)
        return result

    def evolve(self, psi_0, t0, times, iterate=False, n_jobs=1, block_diag=False, stack_state=False,
               imag_time=False, tol=1e-12, krylov_dim=30, dtype=np.complex128, device=False):
        """evolve(::BlockOps, psi_0, t0, times; ...) -- tools.jl:2588-2646.  Returns Npar x n_t
        (or Npar for a scalar time).  n_jobs and block_diag are accepted for API parity: every
        block always runs in the one batched engine.  device=True returns the states as a
        device-resident torch tensor (Npar x n_t view) instead of copying them to the host."""
        if n_jobs <= 0:
            raise ArgumentError("n_jobs must be positive")
        scalar = isinstance(times, numbers.Number)
        targets = np.atleast_1d(np.asarray(times, dtype=np.float64))
        offsets = targets - float(t0)
        scales = -offsets.astype(np.complex128) if imag_time else -1j * offsets
        entries = self._entries(psi_0, dtype)
        res = self._evolution(entries, scales, self._parent.Ns, dtype, tol, krylov_dim)
        if device:
            out = res.T
            if iterate:
                return (out[:, j] for j in range(out.shape[1]))
            return out[:, 0] if scalar else out
        out = res.cpu().numpy().T
        if iterate:
            return (out[:, j].copy() for j in range(out.shape[1]))
        return out[:, 0].copy() if scalar else out


def block_expm(operator, psi_0, H_time_eval=0.0, iterate=False, n_jobs=1, block_diag=False, a=-1j, start=None,
               stop=None, endpoint=True, num=50, shift=None, dtype=np.complex128):
    """block_expm(::BlockOps, psi_0; a, start, stop, num, endpoint) -- tools.jl:2656-2702."""
    if n_jobs <= 0:
        raise ArgumentError("n_jobs must be positive")
    if shift is not None:
        raise ArgumentError("shift is not supported by the GPU block engine")
    if start is None and stop is None:
        grid = np.array([1.0])
    else:
        grid = np.linspace(start, stop, num + (0 if endpoint else 1))
        if not endpoint:
            grid = grid[:-1]
    scales = grid.astype(np.complex128) * complex(a)
    entries = operator._entries(psi_0, dtype)
    res = operator._evolution(entries, scales, operator._parent.Ns, dtype, 1e-12, 30)
    out = res.cpu().numpy().T
    if iterate:
        return (out[:, j].copy() for j in range(out.shape[1]))
    return out[:, 0].copy() if grid.size == 1 else out


def blocks_evolve(ops, kinds, psi_0, scales, tol=1e-12, krylov_dim=30, dtype=np.complex128):
    """One call of qs_blocks_evolve (host vectors): project psi_0 into every block, evolve the active
    blocks in one lockstep batch, recombine -- the C-ABI form of evolve(::BlockOps) for callers without
    device arrays (tools.jl:2445-2646).  Returns parent size x len(scales)."""
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.complex128), np.dtype(np.complex64)):
        raise ArgumentError("block evolution needs a complex dtype (complex128 or complex64)")
    psi = np.ascontiguousarray(np.asarray(psi_0).astype(dt))
    sc = np.ascontiguousarray(np.atleast_1d(np.asarray(scales, dtype=np.complex128)))
    out = np.zeros((psi.size, sc.size), dtype=dt, order="F")
    n = len(ops)
    nmv = np.zeros(max(1, n), dtype=np.int32)
    check(_lib.lib().qs_blocks_evolve((ctypes.c_void_p * n)(*[op._h for op in ops]),
                                      np.asarray(kinds, dtype=np.int32).ctypes.data, n, _tdtype(dt)[1],
                                      psi.ctypes.data, sc.size, sc.ctypes.data, float(tol), int(krylov_dim),
                                      out.ctypes.data, nmv.ctypes.data))
    blocks_evolve.last_n_matvec = nmv[:n].tolist()
    return out
