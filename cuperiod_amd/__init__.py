"""MI355X-native Box Least Squares periodogram engine.

A drop-in for cuPeriod's BLS path: the same ``periodogram`` / ``batch_periodograms``
entry points, ``LightCurve`` / ``GridSpec`` / ``BLSSettings`` inputs and
``Periodogram`` / ``Peak`` / ``BLSPower`` results, computed by hand-written HIP kernels
for gfx950 behind the C ABI of ``include/cuperiod_bls.h`` (backend name ``"hip"``).
"""

from cuperiod_amd.api import best_periods, periodogram, to_input
from cuperiod_amd.batch.runner import BatchSummary, batch_periodograms
from cuperiod_amd.core.columns import Domain
from cuperiod_amd.core.config import BLSSettings, TLSSettings
from cuperiod_amd.core.errors import (
    BackendUnavailableError,
    CuPeriodError,
    DeviceError,
    GridSizeWarning,
    GridTooLargeError,
    InsufficientDataError,
    UnknownMethodError,
)
from cuperiod_amd.core.grid import GridSpec, log_period_grid
from cuperiod_amd.core.lightcurve import LightCurve, MultiBandLightCurve
from cuperiod_amd.core.result import MultiResult, Peak, Periodogram
from cuperiod_amd.methods import bls as _bls  # noqa: F401
from cuperiod_amd.methods import tls as _tls  # noqa: F401
from cuperiod_amd.methods._bls_core import BLSPower, bls_power, bls_power_segments
from cuperiod_amd.methods.base import get_method, list_methods, method_names
from cuperiod_amd.methods.tls import tls_power

__version__ = "0.1.0"

__all__ = [
    "BLSPower",
    "BLSSettings",
    "BackendUnavailableError",
    "BatchSummary",
    "CuPeriodError",
    "DeviceError",
    "Domain",
    "GridSizeWarning",
    "GridSpec",
    "GridTooLargeError",
    "InsufficientDataError",
    "LightCurve",
    "MultiBandLightCurve",
    "MultiResult",
    "Peak",
    "Periodogram",
    "TLSSettings",
    "UnknownMethodError",
    "batch_periodograms",
    "best_periods",
    "bls_power",
    "bls_power_segments",
    "get_method",
    "list_methods",
    "log_period_grid",
    "method_names",
    "periodogram",
    "tls_power",
    "to_input",
]
