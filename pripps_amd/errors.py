"""Error types of the Debye engine.

The variants and their Python exception bases follow the reference:
``Error`` (``src/error.rs:16-74``) and the pyo3 mapping ``into_py_err``
(``pripps-py/src/lib.rs:805-818``): File/Io -> ``IOError``; Yaml,
UnknownAtom, UndefinedFormFactor, QOutOfRange, Validation, Usage ->
``ValueError``; Internal -> ``RuntimeError``.  Engine-only variants
for the device: ``DeviceError`` (a HIP call failed) and
``OutOfDeviceMemoryError`` (admission or allocation refused), both
``RuntimeError``s.  ``NotReadyError`` (a ``TimeoutError``) reports a served
request that was still running when a wait with a time limit ran out.

The integer codes are the C ABI's ``pripps_status`` values.
"""
from __future__ import annotations


class PrippsError(Exception):
    code = 9


class FileError(PrippsError, IOError):
    code = 1


class IoError(PrippsError, IOError):
    code = 2


class YamlError(PrippsError, ValueError):
    code = 3


class UnknownAtomError(PrippsError, ValueError):
    code = 4

    def __init__(self, name: str):
        super().__init__(f"unknown atom name: {name}")


class UndefinedFormFactorError(PrippsError, ValueError):
    code = 5

    def __init__(self, name: str):
        super().__init__(f"undefined form factor: {name}")


class QOutOfRangeError(PrippsError, ValueError):
    code = 6

    def __init__(self, q: float, lo: float, hi: float):
        super().__init__(f"q value {q} outside form factor range [{lo}, {hi}]")
        self.q, self.lo, self.hi = q, lo, hi


class ValidationError(PrippsError, ValueError):
    code = 7


class UsageError(PrippsError, ValueError):
    code = 8


class InternalError(PrippsError, RuntimeError):
    code = 9


class DeviceError(PrippsError, RuntimeError):
    code = 10


class OutOfDeviceMemoryError(PrippsError, RuntimeError):
    code = 11


class NotReadyError(PrippsError, TimeoutError):
    """A served request was still running when the wait timed out."""
    code = 12


_BY_CODE = {c.code: c for c in (FileError, IoError, YamlError, ValidationError, UsageError,
                                InternalError, DeviceError, OutOfDeviceMemoryError,
                                NotReadyError)}


def from_status(code: int, message: str) -> PrippsError:
    """Exception for a non-zero C-ABI status code."""
    if code == 4:
        e = UnknownAtomError.__new__(UnknownAtomError)
        Exception.__init__(e, message)
        return e
    if code == 5:
        e = UndefinedFormFactorError.__new__(UndefinedFormFactorError)
        Exception.__init__(e, message)
        return e
    if code == 6:
        e = QOutOfRangeError.__new__(QOutOfRangeError)
        Exception.__init__(e, message)
        return e
    cls = _BY_CODE.get(code, InternalError)
    return cls(message)
