"""Method interface, registry and settings selection.

The plugin surface of cuPeriod (src/cuperiod/methods/base.py:99-535): a method is a
:class:`PeriodogramMethod` registered under a normalised name; callers resolve a
backend selector (``"auto"``, ``"gpu"``, a concrete name) through
:meth:`PeriodogramMethod.resolve_backend_for`, build a grid with ``default_grid`` and
compute with ``power``. This package registers one method, ``"BLS"``, whose only
backend is ``"hip"`` (the MI355X engine). Selectors that name a CPU or another vendor's
backend raise :class:`BackendUnavailableError`: there is no CPU fallback here.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from collections.abc import Mapping, Sequence
from dataclasses import dataclass
from typing import Any, ClassVar

from cuperiod_amd.core.errors import (
    BackendUnavailableError,
    InsufficientDataError,
    UnknownMethodError,
)
from cuperiod_amd.core.grid import GridSpec
from cuperiod_amd.core.lightcurve import LightCurve
from cuperiod_amd.core.result import Periodogram

#: Generic backend selectors (src/cuperiod/core/backend.py:242-337).
BACKEND_SELECTORS = ("auto", "cpu", "gpu")


def normalize_backend(backend: str) -> str:
    """Selectors ignore case and surrounding whitespace."""
    return str(backend).strip().lower()


def requested_backend(requested: str | None, settings: Any) -> str:
    """``None``/``"auto"`` defer to ``settings.backend``; anything else wins."""
    if requested is None or normalize_backend(requested) == "auto":
        pinned = getattr(settings, "backend", "auto") if settings is not None else "auto"
        return normalize_backend(pinned or "auto")
    return normalize_backend(requested)


def hip_ready() -> bool:
    """Whether the HIP library loads and sees a device."""
    from cuperiod_amd import _lib

    return _lib.device_count() > 0


def searchable_points(lc: LightCurve, min_detections: int, label: str) -> LightCurve:
    """The finite points of ``lc``, refused (``InsufficientDataError``, messages prefixed
    with the method's ``label``) when fewer than ``min_detections`` remain or they span
    no time."""
    finite = lc.finite()
    if finite.n < min_detections:
        raise InsufficientDataError(
            f"{label}: {finite.n} finite points < min_detections {min_detections}")
    if not finite.baseline > 0.0:
        raise InsufficientDataError(f"{label}: no usable time baseline")
    return finite


class PeriodogramMethod(ABC):
    """Abstract periodogram method (same class attributes as cuPeriod's)."""

    name: ClassVar[str]
    objective_sense: ClassVar[str]
    supports_multiband: ClassVar[bool] = False
    natural_domain: ClassVar[Any] = None
    settings_cls: ClassVar[type]
    cpu_backend: ClassVar[str | None] = None
    fast_cpu_backend: ClassVar[str | None] = None
    gpu_backend: ClassVar[str | None] = None
    portable_gpu_backend: ClassVar[str | None] = None
    all_backends: ClassVar[tuple[str, ...]] = ()

    def coerce_settings(self, settings: Any | None) -> Any:
        """A settings instance of this method's type (defaults if ``None``)."""
        wanted = self.settings_cls
        if settings is not None and not isinstance(settings, wanted):
            raise TypeError(f"{self.name} expects {wanted.__name__}, "
                            f"got {type(settings).__name__}")
        return wanted() if settings is None else settings

    def resolve_backend(self, requested: str) -> str:
        """``"auto"``/``"gpu"``/the GPU backend's name -> the GPU backend when the HIP
        library loads and a device is present; anything else raises."""
        requested = normalize_backend(requested)
        if requested in ("auto", "gpu") or requested == self.gpu_backend:
            if self.gpu_backend is not None and hip_ready():
                return self.gpu_backend
            raise BackendUnavailableError(
                f"{self.name}: no HIP device available for backend "
                f"{self.gpu_backend!r} (needs the built gfx950 library and an AMD GPU)"
            )
        if requested == "cpu" or requested in ("numba", "numpy", "astropy"):
            raise BackendUnavailableError(
                f"{self.name}: this engine has no CPU backend ({requested!r}); it runs "
                f"on the GPU as backend {self.gpu_backend!r}"
            )
        if requested in ("cupy", "torch") or requested.startswith("torch:"):
            raise BackendUnavailableError(
                f"{self.name}: backend {requested!r} is not part of this engine; use "
                f"{self.gpu_backend!r}"
            )
        raise BackendUnavailableError(
            f"{self.name}: unknown backend {requested!r}; choose "
            f"{', '.join(repr(s) for s in BACKEND_SELECTORS)} or one of this method's "
            f"backends {self.all_backends}"
        )

    def resolve_backend_for(self, requested: str | None, settings: Any) -> str:
        """Resolve a run's backend: the explicit request, else ``settings.backend``."""
        return self.resolve_backend(requested_backend(requested, settings))

    def is_gpu_backend(self, backend: str, settings: Any = None) -> bool:
        return normalize_backend(backend) == self.gpu_backend

    @abstractmethod
    def default_grid(self, lc: LightCurve, settings: Any) -> GridSpec:
        """The default trial grid of ``lc``."""

    @abstractmethod
    def power(
        self,
        grid: GridSpec,
        lc: LightCurve,
        settings: Any,
        backend: str,
        engine: object | None = None,
    ) -> Periodogram:
        """The periodogram of ``lc`` on ``grid``."""

    def multiband_power(
        self,
        grid: GridSpec,
        mblc: Any,
        settings: Any,
        backend: str,
        engine: object | None = None,
    ) -> Periodogram:
        """The multi-band periodogram of ``mblc`` on ``grid`` (methods with
        ``supports_multiband``; src/cuperiod/methods/base.py:344-357)."""
        raise NotImplementedError(f"{self.name} does not support multi-band input")

    def make_engine(self, backend: str, settings: Any) -> object | None:
        return None

    def estimate_device_bytes(self, n_points: int) -> int:
        return 64 * 1024**2 + n_points * 16 * 6


_REGISTRY: dict[str, PeriodogramMethod] = {}


@dataclass(frozen=True)
class MethodInfo:
    name: str
    objective_sense: str
    supports_multiband: bool
    natural_domain: str
    available_backends: tuple[str, ...]
    all_backends: tuple[str, ...]


def _normalize_name(name: str) -> str:
    return "".join(ch for ch in name if ch.isalnum()).upper()


def register(method: PeriodogramMethod) -> PeriodogramMethod:
    _REGISTRY[_normalize_name(method.name)] = method
    return method


def get_method(name: str) -> PeriodogramMethod:
    """A registered method by name (case and separators ignored)."""
    found = _REGISTRY.get(_normalize_name(name))
    if found is None:
        raise UnknownMethodError(f"unknown method {name!r}; registered: {sorted(_REGISTRY)}")
    return found


def method_names() -> list[str]:
    return sorted(_REGISTRY)


def _info(name: str, method: PeriodogramMethod, ready: bool) -> MethodInfo:
    domain = method.natural_domain
    return MethodInfo(name=name, objective_sense=method.objective_sense,
                      supports_multiband=method.supports_multiband,
                      natural_domain=str(domain) if domain else "any",
                      available_backends=method.all_backends if ready else (),
                      all_backends=method.all_backends)


def list_methods() -> list[MethodInfo]:
    ready = hip_ready()
    return [_info(name, _REGISTRY[name], ready) for name in sorted(_REGISTRY)]


def select_settings(
    methods: Sequence[PeriodogramMethod], settings: Any
) -> dict[str, Any]:
    """Each method's settings from a ``settings=`` argument, by method name: ``None``
    gives every method its defaults; one settings object goes to the methods of its
    class (the others get defaults) and must fit at least one of them; a mapping names
    methods (in any spelling :func:`get_method` accepts) and gives each its settings
    (src/cuperiod/methods/base.py:459-535)."""
    given: dict[str, Any] = {}
    if isinstance(settings, Mapping):
        spelled: dict[str, Any] = {}
        for key, value in settings.items():
            name = get_method(str(key)).name
            if name in spelled:
                raise ValueError(f"settings: {spelled[name]!r} and {key!r} are the same "
                                 f"method ({name}); give it once")
            spelled[name] = key
            given[name] = value
    elif settings is not None:
        fits = [m.name for m in methods if isinstance(settings, m.settings_cls)]
        if not fits:
            wanted = sorted({m.settings_cls.__name__ for m in methods})
            raise TypeError(f"settings: a {type(settings).__name__} fits none of "
                            f"{[m.name for m in methods]} (they take {' or '.join(wanted)}); "
                            "to tune several methods pass {method: settings}")
        given = dict.fromkeys(fits, settings)
    return {m.name: m.coerce_settings(given.get(m.name)) for m in methods}


__all__ = [
    "BACKEND_SELECTORS",
    "MethodInfo",
    "PeriodogramMethod",
    "get_method",
    "hip_ready",
    "list_methods",
    "method_names",
    "normalize_backend",
    "register",
    "requested_backend",
    "select_settings",
]
