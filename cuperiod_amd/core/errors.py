"""Error types of the BLS engine.

The same names and built-in bases as cuPeriod's hierarchy
(src/cuperiod/core/errors.py:33-65, src/cuperiod/core/grid.py:90-102), so callers that
catch ``ValueError``/``RuntimeError`` or the package base class keep working when this
engine is swapped in.
"""

from __future__ import annotations


class CuPeriodError(Exception):
    """Base class of every error this package raises."""


class BackendUnavailableError(CuPeriodError, RuntimeError):
    """The requested backend cannot run here (no HIP library, no device, or a resource
    limit such as the per-workgroup LDS budget was exceeded)."""


class UnknownMethodError(CuPeriodError, ValueError):
    """A periodogram method name that is not registered."""


class InsufficientDataError(CuPeriodError, ValueError):
    """Too few finite points, no usable baseline, an empty grid, or an error column
    with no usable weight."""


class GridTooLargeError(CuPeriodError, ValueError):
    """A trial grid above the size limit; raised before anything is allocated."""


class GridSizeWarning(UserWarning):
    """An automatically chosen trial grid is unusually large (still computed)."""


class DeviceError(CuPeriodError, RuntimeError):
    """The HIP runtime reported a failure (launch, copy or allocation)."""


__all__ = [
    "BackendUnavailableError",
    "CuPeriodError",
    "DeviceError",
    "GridSizeWarning",
    "GridTooLargeError",
    "InsufficientDataError",
    "UnknownMethodError",
]
