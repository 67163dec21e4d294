"""Exception classes named after the Julia exceptions QuSpin.jl throws."""
from . import _lib


class ArgumentError(ValueError):
    """Julia ArgumentError (invalid blocks, operators, keywords)."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch (operator / vector sizes)."""


class ErrorException(RuntimeError):
    """Julia ErrorException (e.g. Krylov exponential did not converge)."""


class DeviceError(RuntimeError):
    """HIP runtime or allocation failure inside libqsgpu."""


def check(status):
    """Raise the exception class mapped from a libqsgpu status code."""
    if status == _lib.QS_OK:
        return
    msg = _lib.last_error()
    if status == _lib.QS_EARG:
        raise ArgumentError(msg)
    if status == _lib.QS_EDIM:
        raise DimensionMismatch(msg)
    if status == _lib.QS_ENOCONV:
        raise ErrorException(msg)
    raise DeviceError("libqsgpu status %d: %s" % (status, msg))
