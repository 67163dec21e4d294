"""ctypes binding of libqsgpu.so (the C ABI declared in include/qsgpu.h)."""
import ctypes
import os
import subprocess

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("QSGPU_LIB") or os.path.join(_PKG_ROOT, "_build", "libqsgpu.so")
CSRC = os.path.join(_PKG_ROOT, "csrc")

QS_OK, QS_EARG, QS_EDIM, QS_ENOCONV, QS_EOOM, QS_EDEVICE = range(6)
QS_UNSET = -1000000
QS_F64, QS_C128, QS_F32, QS_C64 = range(4)
WHICH = {"SA": 0, "LA": 1, "SM": 2, "LM": 3, "BE": 4}
QS_CHECK_SYMM, QS_CHECK_HERM, QS_CHECK_PCON, QS_FORCE_COMPLEX = 1, 2, 4, 8
VARIANT_HOP = 16   # QS_VARIANT_HOP: the two-bit hop matvec kernel
VARIANT_SPLIT = 32   # QS_VARIANT_SPLIT: the split (high word / low word) hop kernel
VARIANT_SELL = 48    # QS_VARIANT_SELL: plain products from the assembled sector matrix
QS_BASIS_EXPLICIT = 1   # qs_basis_create_general: the wide (four-word, stored) form at any L

EXPORTS = (
    "qs_last_error", "qs_device_info", "qs_basis_create", "qs_basis_destroy", "qs_basis_size",
    "qs_basis_states", "qs_basis_reps", "qs_basis_norms", "qs_basis_is_complex", "qs_basis_lookup",
    "qs_op_create", "qs_op_destroy", "qs_op_info", "qs_op_set_diagonal", "qs_op_set_stream",
    "qs_op_stream", "qs_op_matvec", "qs_eigsh", "qs_expm_multiply", "qs_op_set_variant", "qs_op_tune",
    "qs_basis_parent_size", "qs_basis_project", "qs_basis_expand", "qs_batch_create", "qs_batch_destroy",
    "qs_batch_matvec", "qs_batch_expm_multiply", "qs_server_create", "qs_server_destroy", "qs_server_submit",
    "qs_server_wait", "qs_lanczos", "qs_blocks_evolve", "qs_op_assemble", "qs_op_csr", "qs_op_set_graphs",
    "qs_op_matvec_noalloc", "qs_evolve_dynamic", "qs_block_gram", "qs_basis_create_general",
    "qs_basis_states_wide", "qs_basis_lookup_wide",
)
NOALLOC_SCRATCH_BYTES = 32 << 20   # QS_NOALLOC_SCRATCH_BYTES

_lib = None


def build(jobs=8):
    """Compile libqsgpu.so for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", CSRC], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libqsgpu.so is not built (%s); run qsgpu.build() or __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    pp = ctypes.POINTER(ctypes.c_void_p)
    L.qs_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.qs_last_error.restype = ctypes.c_size_t
    L.qs_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, vp, vp]
    L.qs_basis_create.argtypes = [i32] * 11 + [pp]
    L.qs_basis_create_general.argtypes = [i32, i32, i32, i32, vp, vp, i32, pp]
    L.qs_basis_destroy.argtypes = [vp]
    L.qs_basis_size.argtypes = [vp, vp]
    L.qs_basis_states.argtypes = [vp, vp]
    L.qs_basis_states_wide.argtypes = [vp, i32, vp]
    L.qs_basis_lookup_wide.argtypes = [vp, vp, i64, vp, vp]
    L.qs_basis_reps.argtypes = [vp, vp]
    L.qs_basis_norms.argtypes = [vp, vp, vp]
    L.qs_basis_is_complex.argtypes = [vp, vp]
    L.qs_basis_lookup.argtypes = [vp, vp, i64, vp, vp]
    L.qs_op_create.argtypes = [vp, i32, vp, vp, vp, vp, vp, i32, pp]
    L.qs_op_destroy.argtypes = [vp]
    L.qs_op_info.argtypes = [vp, vp, vp, vp]
    L.qs_op_set_diagonal.argtypes = [vp, vp]
    L.qs_op_set_stream.argtypes = [vp, vp]
    L.qs_op_stream.argtypes = [vp, pp]
    L.qs_op_matvec.argtypes = [vp, i32, vp, vp, i64, dbl, dbl, dbl, dbl, i32, i32, i32]
    L.qs_op_matvec_noalloc.argtypes = [vp, i32, vp, vp, i64, dbl, dbl, dbl, dbl, i32, i32]
    L.qs_evolve_dynamic.argtypes = [vp, i32, vp, vp, vp, i32, vp, i64, dbl, i64, vp, i32, dbl, dbl, dbl, vp, vp, vp]
    L.qs_eigsh.argtypes = [vp, i32, i32, i32, dbl, i32, i32, vp, vp, vp, vp, vp]
    L.qs_block_gram.argtypes = [i32, vp, i64, i32, vp, i32, i32, vp]
    L.qs_expm_multiply.argtypes = [vp, i32, vp, i64, vp, dbl, i32, vp, vp]
    L.qs_op_set_variant.argtypes = [vp, i32]
    L.qs_op_tune.argtypes = [vp, i32, i32, vp, vp]
    L.qs_op_assemble.argtypes = [vp]
    L.qs_op_set_graphs.argtypes = [vp, i32]
    L.qs_op_csr.argtypes = [vp, vp, vp, vp, vp]
    L.qs_basis_parent_size.argtypes = [vp, i32, vp]
    L.qs_basis_project.argtypes = [vp, i32, i32, vp, vp]
    L.qs_basis_expand.argtypes = [vp, i32, i32, vp, vp, i64, i32]
    L.qs_batch_create.argtypes = [vp, i32, i32, pp]
    L.qs_batch_destroy.argtypes = [vp]
    L.qs_batch_matvec.argtypes = [vp, i32, vp, vp, vp]
    L.qs_batch_expm_multiply.argtypes = [vp, i32, vp, vp, vp, vp, dbl, i32, vp, vp]
    L.qs_lanczos.argtypes = [vp, i32, vp, i32, i32, dbl, vp, vp, vp, vp, vp, vp]
    L.qs_blocks_evolve.argtypes = [vp, vp, i32, i32, vp, i64, vp, dbl, i32, vp, vp]
    for name in EXPORTS:
        if name != "qs_last_error":
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().qs_last_error(buf, 1024)
    return buf.value.decode(errors="replace")
