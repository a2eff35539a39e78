"""ctypes binding of ``libcuperiod_bls.so`` (the C ABI of ``include/cuperiod_bls.h``).

The library is built in-tree for gfx950 (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C cuperiod_amd/csrc``). There is no fallback: if the library is
missing, or no HIP device is visible, every entry point raises
:class:`~cuperiod_amd.core.errors.BackendUnavailableError`.

Device buffers are torch tensors on the HIP device (torch is used for allocation and
copies only); their raw pointers and the current torch stream go through the ABI.
"""

from __future__ import annotations

import contextlib
import ctypes
import threading
from pathlib import Path

from cuperiod_amd.core.errors import (
    BackendUnavailableError,
    DeviceError,
)

LIB_NAME = "libcuperiod_bls.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

CPB_OK = 0
CPB_E_INVALID = 1
CPB_E_DEVICE = 2
CPB_E_RESOURCE = 3
CPB_E_NODEVICE = 4
N_OUTPUTS = 7
N_JOINT_OUTPUTS = 6
N_TLS_OUTPUTS = 4

#: Every symbol the header declares, in declaration order.
EXPORTED_SYMBOLS = (
    "cpb_version",
    "cpb_device_count",
    "cpb_create",
    "cpb_destroy",
    "cpb_last_error",
    "cpb_status_string",
    "cpb_bls_search",
    "cpb_bls_search_batch",
    "cpb_bls_search_batch_static",
    "cpb_bls_summarize",
    "cpb_peak_candidates",
    "cpb_bls_joint_search",
    "cpb_set_joint_kernel",
    "cpb_tls_search",
    "cpb_linspace_grid",
    "cpb_spectrum_candidates",
    "cpb_select_diverse_peaks",
    "cpb_set_timing",
    "cpb_main_kernel_ms",
    "cpb_set_accumulation",
    "cpb_set_peak_rounding",
    "cpb_set_launch",
    "cpb_graph_create",
    "cpb_graph_launch",
    "cpb_graph_destroy",
    "cpb_synchronize",
)

#: Every symbol the proposer header (include/cuperiod_proposer.h) declares.
PROPOSER_SYMBOLS = (
    "cpp_features",
    "cpp_stack_features",
    "cpp_layer1_forward",
    "cpp_conv_forward",
    "cpp_conv_wgrad",
    "cpp_layer1_wgrad",
    "cpp_flip_weights",
    "cpp_head_train",
    "cpp_head_windows",
    "cpp_adam",
)


class Segment(ctypes.Structure):
    """``cpb_segment``: one duration-homogeneous period block."""

    _fields_ = [
        ("bin_duration", ctypes.c_double),
        ("period_offset", ctypes.c_int64),
        ("n_periods", ctypes.c_int64),
        ("out_offset", ctypes.c_int64),
        ("width", ctypes.c_int32),
        ("dur_offset", ctypes.c_int32),
        ("n_dur", ctypes.c_int32),
        ("curve", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("group", ctypes.c_int32),
        ("lin_start", ctypes.c_double),
        ("lin_step", ctypes.c_double),
        ("lin_stop", ctypes.c_double),
        ("lin_num", ctypes.c_int64),
        ("lin_k0", ctypes.c_int64),
    ]


class Curve(ctypes.Structure):
    """``cpb_curve``: one light curve of a batched search."""

    _fields_ = [
        ("point_offset", ctypes.c_int64),
        ("n_points", ctypes.c_int64),
        ("t_min", ctypes.c_double),
        ("ivar_floor", ctypes.c_double),
        ("sum_y", ctypes.c_double),
        ("sum_ivar", ctypes.c_double),
    ]


class Band(ctypes.Structure):
    """``cpb_band``: one band of a multi-band star."""

    _fields_ = [
        ("point_offset", ctypes.c_int64),
        ("n_points", ctypes.c_int64),
        ("ivar_floor", ctypes.c_double),
    ]


class Star(ctypes.Structure):
    """``cpb_star``: the bands of one star and their common time origin."""

    _fields_ = [
        ("band_offset", ctypes.c_int32),
        ("n_bands", ctypes.c_int32),
        ("t_min", ctypes.c_double),
    ]


class TlsCurve(ctypes.Structure):
    """``cpb_tls_curve``: one light curve of a TLS search."""

    _fields_ = [
        ("point_offset", ctypes.c_int64),
        ("n_points", ctypes.c_int64),
        ("period_offset", ctypes.c_int64),
        ("n_periods", ctypes.c_int64),
        ("out_offset", ctypes.c_int64),
        ("w_total", ctypes.c_double),
    ]


_LOCK = threading.Lock()
_LIB: ctypes.CDLL | None = None
_LOAD_ERROR: str | None = None


def _declare(lib: ctypes.CDLL) -> None:
    vp = ctypes.c_void_p
    i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    lib.cpb_version.restype = ctypes.c_int
    lib.cpb_version.argtypes = []
    lib.cpb_device_count.restype = ctypes.c_int
    lib.cpb_device_count.argtypes = []
    lib.cpb_create.restype = ctypes.c_int
    lib.cpb_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.cpb_destroy.restype = None
    lib.cpb_destroy.argtypes = [vp]
    lib.cpb_last_error.restype = ctypes.c_char_p
    lib.cpb_last_error.argtypes = [vp]
    lib.cpb_status_string.restype = ctypes.c_char_p
    lib.cpb_status_string.argtypes = [ctypes.c_int]
    lib.cpb_bls_search.restype = ctypes.c_int
    lib.cpb_bls_search.argtypes = [
        vp, vp, vp, vp, vp, i64, f64, f64, f64, f64, vp, vp,
        ctypes.POINTER(Segment), i32, i32, i32, vp, i64,
    ]
    lib.cpb_bls_search_batch.restype = ctypes.c_int
    lib.cpb_bls_search_batch.argtypes = [
        vp, vp, vp, vp, vp, i64, ctypes.POINTER(Curve), i32, vp, vp,
        ctypes.POINTER(Segment), i32, i32, i32, vp, i64,
    ]
    lib.cpb_bls_search_batch_static.restype = ctypes.c_int
    lib.cpb_bls_search_batch_static.argtypes = lib.cpb_bls_search_batch.argtypes
    lib.cpb_bls_summarize.restype = ctypes.c_int
    lib.cpb_bls_summarize.argtypes = [vp, vp, vp, vp, vp, i32, vp]
    lib.cpb_peak_candidates.restype = ctypes.c_int
    lib.cpb_peak_candidates.argtypes = [vp, vp, vp, i64, vp, vp, i32, i32, i32, vp, vp, vp]
    lib.cpb_bls_joint_search.restype = ctypes.c_int
    lib.cpb_bls_joint_search.argtypes = [
        vp, vp, vp, vp, vp, i64, ctypes.POINTER(Band), i32, ctypes.POINTER(Star), i32, vp,
        vp, ctypes.POINTER(Segment), i32, i32, i32, vp, i64,
    ]
    lib.cpb_set_joint_kernel.restype = ctypes.c_int
    lib.cpb_set_joint_kernel.argtypes = [vp, ctypes.c_int]
    lib.cpb_tls_search.restype = ctypes.c_int
    lib.cpb_tls_search.argtypes = [
        vp, vp, vp, vp, vp, i64, ctypes.POINTER(TlsCurve), i32, vp, i32,
        ctypes.POINTER(ctypes.c_int32), i32, ctypes.POINTER(f64), f64, vp, i64,
    ]
    lib.cpb_linspace_grid.restype = ctypes.c_int
    lib.cpb_linspace_grid.argtypes = [f64, f64, i64, ctypes.POINTER(f64)]
    lib.cpb_select_diverse_peaks.restype = ctypes.c_int
    lib.cpb_select_diverse_peaks.argtypes = [vp, vp, i64, i32, f64, f64, vp, i32, vp, i32, vp,
                                             ctypes.POINTER(i32), ctypes.POINTER(i64)]
    lib.cpb_spectrum_candidates.restype = ctypes.c_int
    lib.cpb_spectrum_candidates.argtypes = [vp, i64, vp, vp, ctypes.POINTER(i64)]
    lib.cpb_set_timing.restype = ctypes.c_int
    lib.cpb_set_timing.argtypes = [vp, ctypes.c_int]
    lib.cpb_main_kernel_ms.restype = ctypes.c_int
    lib.cpb_main_kernel_ms.argtypes = [vp, ctypes.POINTER(f64)]
    lib.cpb_set_peak_rounding.restype = ctypes.c_int
    lib.cpb_set_peak_rounding.argtypes = [vp, ctypes.c_int]
    lib.cpb_set_launch.restype = ctypes.c_int
    lib.cpb_set_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.cpb_graph_create.restype = ctypes.c_int
    lib.cpb_graph_create.argtypes = [
        vp, vp, vp, vp, i64, ctypes.POINTER(Curve), i32, vp, vp,
        ctypes.POINTER(Segment), i32, i32, i32, vp, i64, vp, vp, vp, ctypes.POINTER(vp),
    ]
    lib.cpb_graph_launch.restype = ctypes.c_int
    lib.cpb_graph_launch.argtypes = [vp, vp, vp]
    lib.cpb_graph_destroy.restype = None
    lib.cpb_graph_destroy.argtypes = [vp]
    lib.cpb_set_accumulation.restype = ctypes.c_int
    lib.cpb_set_accumulation.argtypes = [vp, ctypes.c_int]
    lib.cpb_synchronize.restype = ctypes.c_int
    lib.cpb_synchronize.argtypes = [vp, vp]
    f32 = ctypes.c_float
    proposer = {
        "cpp_features": [vp, vp, vp, vp, vp, vp, vp, vp, i32, f32, vp],
        "cpp_layer1_forward": [vp, vp, vp, vp, i32, vp, vp],
        "cpp_conv_forward": [vp, vp, vp, vp, vp, vp, vp, i32, i32],
        "cpp_conv_wgrad": [vp, vp, vp, i32, vp, vp],
        "cpp_layer1_wgrad": [vp, vp, vp, i32, vp, vp],
        "cpp_flip_weights": [vp, vp, vp],
        "cpp_head_train": [vp, vp, vp, vp, vp, vp, i32, f32, vp, vp, vp, vp],
        "cpp_head_windows": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp],
        "cpp_stack_features": [vp, vp, vp, vp, vp, vp, vp, vp, i32, f32, vp],
        "cpp_adam": [vp, vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, i32],
    }
    for name, args in proposer.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args


def load_library() -> ctypes.CDLL:
    """The loaded library; :class:`BackendUnavailableError` if it cannot be loaded."""
    global _LIB, _LOAD_ERROR
    with _LOCK:
        if _LIB is not None:
            return _LIB
        if not LIB_PATH.exists():
            _LOAD_ERROR = (
                f"BLS: the HIP library {LIB_PATH} is not built; run "
                "`make -C cuperiod_amd/csrc` (hipcc, gfx950)"
            )
            raise BackendUnavailableError(_LOAD_ERROR)
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
        # Loading torch first makes this library bind to that one runtime, so torch's
        # streams and allocations and this library's launches share it; loading this
        # library first would hand torch a runtime it was not built against.
        import torch  # noqa: F401

        try:
            lib = ctypes.CDLL(str(LIB_PATH))
        except OSError as exc:
            _LOAD_ERROR = f"BLS: cannot load {LIB_PATH}: {exc}"
            raise BackendUnavailableError(_LOAD_ERROR) from exc
        _declare(lib)
        _LIB = lib
        return lib


def library_available() -> bool:
    """Whether the HIP library loads (no device needed)."""
    try:
        load_library()
    except BackendUnavailableError:
        return False
    return True


def device_count() -> int:
    """HIP devices the library sees (0 without the library or a device)."""
    try:
        return int(load_library().cpb_device_count())
    except BackendUnavailableError:
        return 0


def status_error(status: int, message: str) -> Exception:
    """The Python exception for a C-ABI status code."""
    if status == CPB_E_INVALID:
        return ValueError(message)
    if status in (CPB_E_RESOURCE, CPB_E_NODEVICE):
        return BackendUnavailableError(message)
    return DeviceError(message)


class Context:
    """A ``cpb_context`` on one device; calls are serialised by a lock, so one
    context may be shared by threads (each call runs alone)."""

    def __init__(self, device: int = 0) -> None:
        self.lib = load_library()
        handle = ctypes.c_void_p()
        status = self.lib.cpb_create(int(device), ctypes.byref(handle))
        if status != CPB_OK:
            reason = self.lib.cpb_status_string(status).decode()
            raise status_error(
                status if status != CPB_E_INVALID else CPB_E_NODEVICE,
                f"BLS: cannot open HIP device {device} ({reason})",
            )
        self.handle = handle
        self.device = int(device)
        self.lock = threading.RLock()

    def check(self, status: int, what: str) -> None:
        if status == CPB_OK:
            return
        detail = self.lib.cpb_last_error(self.handle).decode(errors="replace")
        raise status_error(status, f"BLS {what}: {detail}")

    @contextlib.contextmanager
    def peak_rounding(self, enabled: bool):
        """Within the block, searches on this context do (or do not) search each
        segment's near-maximum periods again with the reference's FP64 sequence
        (``cpb_set_peak_rounding``; on by default)."""
        with self.lock:
            self.check(self.lib.cpb_set_peak_rounding(self.handle, int(bool(enabled))),
                       "peak rounding")
            try:
                yield self
            finally:
                self.check(self.lib.cpb_set_peak_rounding(self.handle, 1), "peak rounding")

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.cpb_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_CONTEXTS: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    """The process-wide context of ``device`` (created on first use)."""
    with _LOCK:
        ctx = _CONTEXTS.get(device)
    if ctx is None:
        ctx = Context(device)
        with _LOCK:
            _CONTEXTS.setdefault(device, ctx)
            ctx = _CONTEXTS[device]
    return ctx


__all__ = [
    "EXPORTED_SYMBOLS",
    "PROPOSER_SYMBOLS",
    "LIB_PATH",
    "N_JOINT_OUTPUTS",
    "N_OUTPUTS",
    "N_TLS_OUTPUTS",
    "Band",
    "Context",
    "Curve",
    "Segment",
    "Star",
    "TlsCurve",
    "context",
    "device_count",
    "library_available",
    "load_library",
    "status_error",
]
