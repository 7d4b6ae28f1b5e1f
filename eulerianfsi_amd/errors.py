"""Exceptions of the device backend, mapped from fsi_status codes to the
exception types the reference throws (src/TimeIntegrationIMEX.jl,
src/Projection.jl, src/StructureTensor.jl)."""


class ArgumentError(ValueError):
    """Julia ArgumentError: invalid value or a failed stage-1 / compatibility check."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch: an array has the wrong shape."""


class SolverError(RuntimeError):
    """Julia ErrorException: projection residual or constraint check failed."""


class DomainError(ValueError):
    """Julia DomainError: non-SPD structure tensor under the geometric extension."""


class UnsupportedError(NotImplementedError):
    """Configuration outside the device path; a caller keeps the CPU path for it."""


class DeviceError(RuntimeError):
    """HIP runtime failure."""


_MAP = {1: ArgumentError, 2: DimensionMismatch, 3: SolverError, 4: DomainError, 5: UnsupportedError, 6: DeviceError}


def raise_for_status(status: int, message: str) -> None:
    raise _MAP.get(status, DeviceError)(message)
