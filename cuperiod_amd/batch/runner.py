"""``batch_periodograms``: BLS over many light curves, one launch per chunk.

cuPeriod's batch API (src/cuperiod/batch/runner.py:151-185, 734-1032)
with the same arguments, ``BatchSummary``, row format, sinks and resume rules, but a
different execution model. cuPeriod's ``_process_chunk`` (``:214-250``) calls
``method.power`` once per light curve, each with its own uploads, launches and a
``(7, F)`` download; a GPU pool of processes hides part of that. Here every chunk of
light curves is packed into one set of device arrays and searched by ONE batched engine
call (:mod:`cuperiod_amd.batch.engine`), in this process; the next chunk is packed on
the host while the device searches the current one. Each curve's rows are the rows that
``periodogram(lc, "BLS")`` gives: the device regenerates each curve's default grid with
numpy's recipe. Rows without the raw spectrum are built from device peak candidates
(``cpb_peak_candidates``: the candidates above a per-curve threshold, a few KB per curve
instead of the (7, F) spectrum); the selection is the host's own greedy on them, and a
curve whose selection they cannot decide (it would read past them, or equal scores make
its order depend on the whole spectrum) is built from its full spectrum. Their ``sde_k``
uses the device's mean and standard deviation of the power (numpy's pairwise summation
differs in the last bits); every other column is identical.

Sinks (``:734-761``): ``None`` returns rows in memory; a ``.parquet``/``.csv`` file is
rewritten at the end (merged with its rows when resuming); a directory gets one
``part-NNNNN.parquet`` per chunk and a ``_manifest.json`` recording the request, so a
resumed run that would write different rows is refused (``:546-693``). Resuming skips
every ``(key, method)`` pair the sink already holds.
"""

from __future__ import annotations

import ctypes
import dataclasses
import hashlib
import itertools
import json
import re
import time
import warnings
from collections.abc import Callable, Mapping, Sequence
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.batch.io import (
    InputItem,
    alias_diverse_for,
    csv_safe,
    peak_extra_columns,
    peaks_to_row,
    periodogram_to_row,
    read_rows,
    resolve_inputs,
    source_label,
    write_rows,
)
from cuperiod_amd.core.columns import Domain, coerce_domain
from cuperiod_amd.core.errors import InsufficientDataError
from cuperiod_amd.core.grid import GridSpec
from cuperiod_amd.core.lightcurve import LightCurve, MultiBandLightCurve, raise_if_unusable
from cuperiod_amd.core.result import Peak, Periodogram, peak_defaults_from_settings
from cuperiod_amd.methods._bls_core import _OUTPUT_NAMES
from cuperiod_amd.methods.base import (
    get_method,
    normalize_backend,
    requested_backend,
    searchable_points,
    select_settings,
)

#: Light curves per chunk (one engine call, one part file of a directory sink).
DEFAULT_CHUNK_SIZE = 256

_PART_RE = re.compile(r"^part-(\d{5,})\.parquet$")
_STALE_PART_RE = re.compile(r"^part-\d{5,}\.parquet\.tmp$")
_PART_TEMP_RE = re.compile(r"^\.part-\d{5,}\.parquet\.\d+\.[0-9a-f]+\.tmp$")

#: Settings fields that choose hardware or batch sizes, not results: left out of a
#: directory sink's fingerprint so a run can resume elsewhere.
_RUNTIME_SETTINGS_FIELDS = frozenset(
    {"backend", "device", "precision", "batch_periods", "accumulation"}
)
_FINGERPRINT_FIELDS = ("n_best", "store_raw", "columns", "domain", "band_column", "grid")


@dataclass
class BatchSummary:
    """Outcome of :func:`batch_periodograms`.

    ``n_done`` rows produced by this run, ``n_failed`` light curves that raised (their
    ``(key, message)`` in ``errors``, input order), ``n_skipped`` light curves already
    complete in a resumed sink; ``rows`` (input order) when ``sink is None``;
    ``n_workers`` processes used (1: this one; 0: nothing to compute).
    """

    n_inputs: int
    n_done: int
    n_failed: int
    n_skipped: int
    methods: tuple[str, ...]
    device: str
    sink: str | None
    errors: list[tuple[str, str]] = field(default_factory=list)
    rows: list[dict[str, Any]] | None = None
    n_workers: int | None = None


@dataclass(frozen=True)
class _ChunkConfig:
    methods: tuple[str, ...]
    backends: Mapping[str, str]
    settings_map: Mapping[str, Any]
    n_best: int
    store_raw: bool
    grid: GridSpec | None
    extra_keys: tuple[str, ...]


# --- backend choice ------------------------------------------------------------------


def _worker_backend(device: str, backend: str | None, settings: Any) -> str:
    """The backend selector for ``device``: ``"cpu"`` asks for a CPU backend (which
    this package does not have, so resolving it raises), ``"gpu"`` maps ``"auto"`` to
    ``"gpu"`` and refuses an explicit CPU backend."""
    if device == "cpu":
        explicit = requested_backend(backend, settings)
        return "cpu" if explicit in ("auto", "gpu") else explicit
    explicit = "auto" if backend is None else normalize_backend(backend)
    if explicit == "auto":
        pinned = requested_backend(None, settings)
        return "gpu" if pinned == "auto" else pinned
    if explicit in ("gpu", "hip"):
        return explicit
    raise ValueError(
        f"backend {explicit!r} runs on the CPU, but device='gpu' asks for GPU workers; "
        "use a GPU backend ('auto', 'gpu' or 'hip')"
    )


# --- sinks -----------------------------------------------------------------------------


_FILE_SINK_SUFFIXES = frozenset({".parquet", ".pq", ".csv"})


def _classify_sink(sink: str | Path | None) -> tuple[str, Path, Path]:
    """``("memory"|"file"|"dir", directory, path)``; a file sink's directory is made."""
    if sink is None:
        return "memory", Path(), Path()
    path = Path(sink)
    suffix = path.suffix.lower()
    is_dir = path.is_dir()
    if suffix not in _FILE_SINK_SUFFIXES:
        if suffix and not is_dir:
            raise ValueError(
                f"unsupported sink {str(sink)!r}: a file sink must end in .parquet or .csv; "
                "give a directory (no extension) for a resumable multi-part sink.")
        return "dir", path, path
    if is_dir:
        raise ValueError(
            f"sink {str(sink)!r} is an existing directory, yet its {suffix} suffix asks "
            "for a single-file sink; give a file path, or a directory path without an "
            "extension for a resumable multi-part sink.")
    path.parent.mkdir(parents=True, exist_ok=True)
    return "file", path.parent, path


def _part_path(sink_dir: Path, idx: int) -> Path:
    return sink_dir / f"part-{idx:05d}.parquet"


def _sink_key(value: Any, *, csv_sink: bool) -> str | None:
    if value is None:
        return None
    text = str(value)
    return str(csv_safe(text)) if csv_sink else text


def _corrupt_errors() -> tuple[type[BaseException], ...]:
    import csv

    import pyarrow as pa

    return (pa.ArrowInvalid, UnicodeDecodeError, csv.Error)


def _quarantine(path: Path, *, hidden: bool) -> Path:
    """Move an unreadable sink file aside (never delete it): to
    ``[.]<name>.corrupt-<timestamp>[-<n>]``, the first such name that is free."""
    stem = f"{'.' if hidden else ''}{path.name}.corrupt-{time.strftime('%Y%m%dT%H%M%S')}"
    candidates = (stem if n == 0 else f"{stem}-{n}" for n in itertools.count())
    target = next(path.with_name(name) for name in candidates
                  if not path.with_name(name).exists())
    path.replace(target)
    return target


def _done_pairs(rows: Sequence[Mapping[str, Any]], *, csv_sink: bool) -> dict:
    done: dict[tuple[str | None, str | None], str] = {}
    for row in rows:
        ident = (_sink_key(row.get("key"), csv_sink=csv_sink),
                 _sink_key(row.get("method"), csv_sink=csv_sink))
        done.setdefault(ident, str(row.get("source") or ""))
    return done


def _file_sink_done(path: Path) -> dict:
    """``(key, method)`` pairs an existing file sink holds; an unreadable file is
    moved aside with a warning and counts as empty."""
    if not path.exists():
        return {}
    try:
        rows = read_rows(path)
    except _corrupt_errors() as exc:
        moved = _quarantine(path, hidden=False)
        warnings.warn(
            f"existing sink {path} is not a readable {path.suffix} file ({exc}); it was "
            f"moved to {moved} and a new file will be written",
            UserWarning,
            stacklevel=3,
        )
        return {}
    return _done_pairs(rows, csv_sink=path.suffix.lower() == ".csv")


def _scan_parts(sink_dir: Path) -> tuple[dict, int]:
    """Pairs a directory sink's parts hold, and the next free part index."""
    done: dict = {}
    next_index = 0
    if not sink_dir.is_dir():
        return done, next_index
    for entry in sorted(sink_dir.iterdir()):
        if _STALE_PART_RE.match(entry.name):
            entry.unlink(missing_ok=True)
            continue
        match = _PART_RE.match(entry.name)
        if match is None:
            continue
        next_index = max(next_index, int(match.group(1)) + 1)
        try:
            rows = read_rows(entry)
        except _corrupt_errors() as exc:
            moved = _quarantine(entry, hidden=True)
            warnings.warn(
                f"part file {entry} is unreadable ({exc}); it was moved to {moved.name} "
                "and its light curves will be computed again",
                UserWarning,
                stacklevel=3,
            )
            continue
        for ident, src in _done_pairs(rows, csv_sink=False).items():
            done.setdefault(ident, src)
    return done, next_index


def _purge_parts(sink_dir: Path) -> None:
    if not sink_dir.is_dir():
        return
    for entry in sink_dir.iterdir():
        name = entry.name
        if (_PART_RE.match(name) or _STALE_PART_RE.match(name)
                or _PART_TEMP_RE.match(name)) and entry.is_file():
            entry.unlink()


def _grid_fingerprint(grid: GridSpec | None) -> dict[str, Any] | None:
    if grid is None:
        return None
    digest = hashlib.sha256(np.ascontiguousarray(grid.values).tobytes()).hexdigest()[:16]
    return {"kind": str(grid.kind), "size": int(grid.values.size), "sha256": digest}


def _run_fingerprint(methods, settings_map, *, n_best, store_raw, columns, domain,
                     band_column, grid, extra_keys) -> dict[str, Any]:
    raw = {
        "n_best": int(n_best),
        "store_raw": bool(store_raw),
        "columns": None if columns is None else str(columns),
        "domain": domain.value if domain is not None else None,
        "band_column": band_column,
        "grid": _grid_fingerprint(grid),
        "extra_keys": list(extra_keys),
        "settings": {
            name: settings_map[name].model_dump(mode="json",
                                                exclude=set(_RUNTIME_SETTINGS_FIELDS))
            for name in methods
        },
    }
    return json.loads(json.dumps(raw))


def _settings_changes(stored: Any, current: Mapping[str, Any]) -> list[str]:
    """Per method whose settings both requests record: the fields that differ."""
    if not isinstance(stored, Mapping):
        return []
    out = []
    for method, now in current.items():
        before = stored.get(method)
        if isinstance(before, Mapping):
            fields = sorted(f for f in set(before) & set(now) if before[f] != now[f])
            if fields:
                out.append(f"{method} settings ({', '.join(fields)})")
    return out


def _merge_config(stored: Mapping[str, Any], current: Mapping[str, Any],
                  sink_dir: Path) -> dict[str, Any]:
    """Refuse a resumed request whose rows would differ from the directory's; a new
    method is recorded."""
    problems = [name for name in _FINGERPRINT_FIELDS
                if name in stored and stored[name] != current.get(name)]
    columns = list(current.get("extra_keys", []))
    if isinstance(stored.get("extra_keys"), list):
        missing = [k for k in columns if k not in stored["extra_keys"]]
        if missing:
            problems.append(f"new per-peak columns {missing} that the existing part files "
                            "lack; write that method to a directory of its own")
        columns = list(stored["extra_keys"])
    problems += _settings_changes(stored.get("settings"), current.get("settings", {}))
    if problems:
        raise ValueError(
            f"directory sink {sink_dir} was written with a different request "
            f"({'; '.join(problems)}): resuming would mix rows that do not belong "
            "together. Repeat the original request, use a fresh directory, or pass "
            "resume=False to replace the directory's part files.")
    before = stored.get("settings") if isinstance(stored.get("settings"), Mapping) else {}
    return {**current, "extra_keys": columns,
            "settings": {**before, **current.get("settings", {})}}


def _dir_manifest_guard(sink_dir: Path, resume: bool, config: Mapping[str, Any]) -> dict:
    """Check and write a directory sink's ``_manifest.json`` before any work."""
    manifest = sink_dir / "_manifest.json"
    stored: dict[str, Any] = {}
    if manifest.exists():
        try:
            loaded = json.loads(manifest.read_text(encoding="utf-8"))
            stored = loaded if isinstance(loaded, dict) else {}
        except Exception:  # noqa: BLE001 - a corrupt manifest must not abort the run
            stored = {}
    if not resume:
        _purge_parts(sink_dir)
    elif sink_dir.is_dir() and any(_PART_RE.match(e.name) for e in sink_dir.iterdir()):
        if isinstance(stored.get("config"), Mapping):
            config = _merge_config(stored["config"], config, sink_dir)
    sink_dir.mkdir(parents=True, exist_ok=True)
    record = {"config": dict(config)}
    manifest.write_text(json.dumps(record, indent=1, sort_keys=True), encoding="utf-8")
    return record


def _finalize_file(rows: list[dict[str, Any]], path: Path, resume: bool) -> None:
    """Write a file sink; when resuming, keep its rows and add only new pairs."""
    if not (resume and path.exists()):
        write_rows(rows, path)
        return
    if not rows:
        return
    try:
        existing = read_rows(path)
    except FileNotFoundError:
        existing = []
    except _corrupt_errors() as exc:
        moved = _quarantine(path, hidden=False)
        warnings.warn(
            f"could not read existing sink {path} ({exc}); it was moved to {moved} and a "
            "new file is written",
            UserWarning,
            stacklevel=3,
        )
        existing = []
    if not existing:
        write_rows(rows, path)
        return
    csv_sink = path.suffix.lower() == ".csv"
    seen = set(_done_pairs(existing, csv_sink=csv_sink))
    new = [r for r in rows if (_sink_key(r.get("key"), csv_sink=csv_sink),
                               _sink_key(r.get("method"), csv_sink=csv_sink)) not in seen]
    if new:
        write_rows(existing + new, path)


# --- the batched search ------------------------------------------------------------


@dataclass
class _Prepared:
    """A light curve that passed the per-curve checks, ready to pack."""

    position: int
    key: str
    label: str
    finite: LightCurve
    periods: np.ndarray | None = None
    # the light curve's own grid when ``batch_periodograms`` was given a grid rule
    grid: Any = None


def _prepare(position: int, key: str, lc: LightCurve, settings: Any,
             grid: GridSpec | None) -> _Prepared:
    """The checks ``method.power`` makes before searching one light curve
    (src/cuperiod/methods/bls.py:371-389), raising the same errors."""
    from cuperiod_amd.batch import engine
    from cuperiod_amd.methods.bls import _grid_segments

    raise_if_unusable(lc, "BLS")
    finite = searchable_points(lc.in_domain(Domain.FLUX), settings.min_detections, "BLS")
    own = None
    if callable(grid):
        # a grid rule: this light curve's grid (a GridSpec or a parametric LinspaceGrid)
        own = grid(finite)
        if not isinstance(own, (GridSpec, engine.LinspaceGrid)):
            raise TypeError("a grid rule must return a GridSpec or a LinspaceGrid, got "
                            f"{type(own).__name__}")
        grid = own if isinstance(own, GridSpec) else None
        if grid is None:
            return _Prepared(position, key, source_label(lc), finite, grid=own)
    if grid is None:
        plan = engine._default_plan(0, finite.baseline, finite.time, settings)
        if not plan.segments:
            raise InsufficientDataError(
                "BLS: baseline too short for the configured period range"
            )
    elif not _grid_segments(grid, settings):
        raise InsufficientDataError("BLS: empty trial grid")
    return _Prepared(position, key, source_label(lc), finite, grid=own)


def _prepare_tls(position: int, key: str, lc: LightCurve, settings: Any,
                 grid: GridSpec | None) -> _Prepared:
    """The checks and the grid of ``TLSMethod.power`` for one light curve
    (src/cuperiod/methods/tls.py:767-797), raising the same errors."""
    from cuperiod_amd.methods.tls import TLSMethod

    raise_if_unusable(lc, "TLS")
    finite = searchable_points(lc.in_domain(Domain.FLUX), settings.min_detections, "TLS")
    g = grid if grid is not None else TLSMethod().default_grid(finite, settings)
    return _Prepared(position, key, source_label(lc), finite,
                     np.asarray(g.period, dtype=np.float64))


def _tls_rows(prepared: Sequence[_Prepared], settings: Any, cfg: "_ChunkConfig"):
    """One TLS launch for the chunk and its rows (the periodograms of
    ``TLSMethod.power``)."""
    from cuperiod_amd.methods.tls import tls_periodogram, tls_power_batch

    found = tls_power_batch(
        [(p.finite.time, p.finite.value, p.finite.error) for p in prepared],
        [p.periods for p in prepared], settings=settings)
    rows = []
    for prep, res in zip(prepared, found):
        pg = tls_periodogram(prep.finite, prep.periods, res, "hip")
        pg = pg.with_peak_defaults(peak_defaults_from_settings(settings))
        rows.append(periodogram_to_row(prep.key, pg, n_best=cfg.n_best,
                                       store_raw=cfg.store_raw, extra_keys=cfg.extra_keys,
                                       source=prep.label))
    return rows


def _multiband_row(key: str, mblc: MultiBandLightCurve, name: str, settings: Any,
                   cfg: "_ChunkConfig") -> dict[str, Any]:
    """The row of one multi-band light curve: ``method.multiband_power`` on the given
    grid or the stacked bands' default grid (src/cuperiod/batch/runner.py:185-207)."""
    from cuperiod_amd.api import _multiband_grid

    method = get_method(name)
    if not method.supports_multiband:
        raise ValueError(f"{method.name} does not support multi-band input")
    raise_if_unusable(mblc, method.name)
    g = cfg.grid if cfg.grid is not None else _multiband_grid(method, mblc, settings)
    pg = method.multiband_power(g, mblc, settings, "hip")
    pg = pg.with_peak_defaults(peak_defaults_from_settings(settings))
    return periodogram_to_row(key, pg, n_best=cfg.n_best, store_raw=cfg.store_raw,
                              extra_keys=cfg.extra_keys, source=source_label(mblc))


def _curve_arrays(lc: LightCurve) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    err = lc.error if lc.error is not None else np.ones_like(lc.value)
    return lc.time, lc.value, err


def _periodogram(prep: _Prepared, periods: np.ndarray, spectrum: Mapping[str, np.ndarray],
                 settings: Any) -> Periodogram:
    """The single-curve assembly (``_assemble``) of one curve's batched spectrum."""
    from cuperiod_amd.methods.bls import _assemble

    chunks = {
        "period": [periods],
        "power": [spectrum["power"]],
        "depth": [spectrum["depth"]],
        "depth_snr": [spectrum["depth_snr"]],
        "duration": [spectrum["duration"]],
        "transit_time": [spectrum["transit_time"]],
    }
    pg = _assemble(chunks, prep.finite.n, prep.finite.baseline, "hip", prep.finite.meta)
    return pg.with_peak_defaults(peak_defaults_from_settings(settings))


def _launch(prepared: Sequence[_Prepared], settings: Any, grid: GridSpec | None) -> Any:
    """Pack the curves and enqueue one batched search (returns without waiting)."""
    from cuperiod_amd.batch import engine

    import torch

    grids = [p.grid if p.grid is not None else grid for p in prepared]
    pb = engine.pack([_curve_arrays(p.finite) for p in prepared], grids, settings)
    out, summary = engine.run_packed(pb)
    searched = torch.cuda.Event()
    searched.record(torch.cuda.current_stream(out.device))
    return pb, out, summary, searched


_FETCH_STREAMS: dict[int, Any] = {}


def _fetch_stream(device) -> Any:
    """A stream of the device's own for reading finished chunks back (default priority),
    so the reads wait for their chunk's search only, not for the search of the next
    chunk queued behind it on torch's current stream."""
    import torch

    stream = _FETCH_STREAMS.get(device.index)
    if stream is None:
        stream = _FETCH_STREAMS[device.index] = torch.cuda.Stream(device)
    return stream


def _collect(handle: Any, *, candidates: bool = False) -> list:
    """Wait for a launch. Per curve: ``(periods, spectrum)`` (ascending period), or,
    with ``candidates``, a :class:`_Candidates` (the rows need no full spectrum)."""
    from cuperiod_amd.batch import engine

    import torch

    pb, out, summary, searched = handle
    if not candidates:
        res = engine.finish(pb, out, summary, spectra=True)
        return [(pb.plans[c].periods(), res.spectra[c]) for c in range(len(pb.plans))]
    side = _fetch_stream(out.device)
    side.wait_event(searched)
    with torch.cuda.stream(side):
        pk = engine.peak_candidates(pb, out)
        summ = summary.cpu().numpy()
    reader = _ChunkReader(out, pb.curve_base, pb.curve_count, side)
    results = []
    for c, plan in enumerate(pb.plans):
        n = int(pb.curve_count[c])
        base = int(pb.curve_base[c])

        def fetch(c=c, plan=plan, n=n, base=base):
            with torch.cuda.stream(side):
                host = out[:, base:base + n].cpu().numpy()
            return plan.periods(), {name: host[i].copy()
                                    for i, name in enumerate(_OUTPUT_NAMES)}

        x = pk.index[c]
        results.append(_Candidates(
            periods=plan.periods_at(n - 1 - x) if x.size else np.zeros(0),
            values=pk.values[c], total=int(pk.total[c]), written=int(pk.written[c]),
            cap=pk.cap, mean=float(summ[c, 0]), std=float(summ[c, 1]), fetch=fetch,
            fetch_power=lambda c=c: reader.power_rows([c])[c],
            fetch_extras=lambda k, c=c: reader.extras([(c, k)])[0],
            periods_at=plan.periods_at, reader=reader, curve=c,
        ))
    return results


#: Outputs a peak's extras are read from (depth, depth_snr, duration, transit_time).
_EXTRA_ROWS = (1, 3, 4, 5)


class _ChunkReader:
    """Reads finished outputs of one chunk back on the fetch stream, many curves per
    transfer: power rows, and the extras at given columns."""

    def __init__(self, out: Any, base: np.ndarray, count: np.ndarray, stream: Any) -> None:
        self.out, self.base, self.count, self.stream = out, base, count, stream

    def power_rows(self, curves: Sequence[int]) -> dict[int, np.ndarray]:
        """Each curve's power row (ascending period)."""
        import torch

        if not curves:
            return {}
        spans = [(int(self.base[c]), int(self.count[c])) for c in curves]
        with torch.cuda.stream(self.stream):
            host = torch.cat([self.out[0, b:b + n] for b, n in spans]).cpu().numpy()
        cuts = np.cumsum([n for _, n in spans])[:-1]
        return dict(zip(curves, np.split(host, cuts)))

    def extras(self, requests: Sequence[tuple[int, np.ndarray]]) -> list[np.ndarray]:
        """For each ``(curve, columns)``: the ``(len(columns), 4)`` extras there."""
        import torch

        sizes = [len(k) for _, k in requests]
        if sum(sizes) == 0:
            return [np.zeros((0, len(_EXTRA_ROWS))) for _ in requests]
        cols = np.concatenate([np.asarray(k, dtype=np.int64) + int(self.base[c])
                               for c, k in requests])
        with torch.cuda.stream(self.stream):
            index = torch.from_numpy(cols).to(self.out.device)
            picked = self.out.index_select(1, index)[list(_EXTRA_ROWS)]
            host = picked.cpu().numpy().T
        return np.split(host, np.cumsum(sizes)[:-1])


@dataclass
class _Candidates:
    """One curve's device peak candidates in ascending frequency, its SDE statistics and
    ways to fetch its full spectrum, its power row, or outputs at given columns."""

    periods: np.ndarray
    values: np.ndarray
    total: int
    written: int
    cap: int
    mean: float
    std: float
    fetch: Callable[[], tuple]
    fetch_power: Callable[[], np.ndarray] | None = None
    fetch_extras: Callable[[np.ndarray], np.ndarray] | None = None
    periods_at: Callable[[np.ndarray], np.ndarray] | None = None
    reader: Any = None
    curve: int = -1


def _peaks_from(period, freq, power, extras, chosen, mean, std) -> list[Peak]:
    peaks = []
    for rank, i in enumerate(chosen, 1):
        p = float(power[i])
        e = extras[rank - 1]
        sde = (p - mean) / std if std > 0.0 else 0.0
        peaks.append(Peak(period=float(period[i]), frequency=float(freq[i]), power=p, rank=rank,
                          extra={"depth": float(e[0]), "depth_snr": float(e[1]),
                                 "duration": float(e[2]), "t0": float(e[3]), "sde": sde}))
    return peaks


def _stub(prep: _Prepared, settings: Any, freq, period, score) -> Periodogram:
    return Periodogram(method="BLS", backend="hip", frequency=freq, period=period,
                       power=score, objective_sense="max", n_samples=prep.finite.n,
                       baseline=prep.finite.baseline,
                       peak_defaults=peak_defaults_from_settings(settings))


def _row_of(prep: _Prepared, cfg: _ChunkConfig, peaks: list[Peak]) -> dict[str, Any]:
    return peaks_to_row(prep.key, "BLS", "hip", prep.finite.n, prep.finite.baseline, peaks,
                        n_best=cfg.n_best, extra_keys=cfg.extra_keys, source=prep.label)


def _row_from_candidates(prep: _Prepared, cand: _Candidates, settings: Any,
                         cfg: _ChunkConfig) -> dict[str, Any] | None:
    """The row when the device candidates decide the selection, else ``None``."""
    if not (cand.written <= cand.cap and cand.periods.size):
        return None
    count = max(int(cfg.n_best), 1)
    freq = 1.0 / cand.periods
    period = 1.0 / freq  # what Periodogram.from_spectrum stores
    score = cand.values[:, 0]
    seen: list[int] = []
    chosen = _stub(prep, settings, freq, period, score).select_peaks(
        freq, score, np.arange(score.size), count, alias_diverse=alias_diverse_for("BLS"),
        examined=seen)
    # Equal scores among the visited candidates (and the one after them) leave their
    # order to how numpy sorts the whole candidate list. That order cannot change the
    # outcome when no chosen peak has such a score: every member of a tied group is then
    # blocked by a peak of higher score, chosen before the group in any order.
    visited = np.sort(score)[::-1][:min(seen[0] + 1, score.size)]
    values, counts = np.unique(visited, return_counts=True)
    tied = values[counts > 1]
    decided = (chosen.size == count or cand.written == cand.total) and (
        tied.size == 0 or not np.isin(score[chosen], tied).any())
    if not decided:
        return None
    return _row_of(prep, cfg, _peaks_from(period, freq, score, cand.values[chosen, 1:],
                                          chosen, cand.mean, cand.std))


def _spectrum_candidates(power: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The single-curve path's peak candidates of an engine power row (ascending period)
    on its ascending-frequency spectrum, non-finite powers read as 0: their indices, in
    the order that path lists them, and their scores (``cpb_spectrum_candidates``)."""
    lib = _lib.load_library()
    row = np.ascontiguousarray(power, dtype=np.float64)
    index = np.empty(row.size // 2 + 2, dtype=np.int64)
    score = np.empty(row.size // 2 + 2, dtype=np.float64)
    count = ctypes.c_int64()
    status = lib.cpb_spectrum_candidates(row.ctypes.data, row.size, index.ctypes.data,
                                         score.ctypes.data, ctypes.byref(count))
    if status != 0:
        raise RuntimeError(f"cpb_spectrum_candidates: status {status}")
    return index[:count.value], score[:count.value]


def _select_on_power(prep: _Prepared, cand: _Candidates, settings: Any, cfg: _ChunkConfig,
                     power: np.ndarray) -> tuple:
    """The single-curve path's selection on the full power row (ascending period):
    ``(period, freq, score, chosen, columns)``, the first three over the row's peak
    candidates (``chosen`` indexes them), ``columns`` the chosen samples' engine columns.
    The candidates' scores are the array the single-curve path sorts (the same values in
    the same order), so equal scores fall into the same order; only the candidates'
    periods are formed."""
    n = int(power.size)
    index, score = _spectrum_candidates(power)
    columns = n - 1 - index
    freq = 1.0 / cand.periods_at(columns)
    period = 1.0 / freq
    pg = _stub(prep, settings, freq, period, score)
    chosen = pg.select_peaks(freq, score, np.arange(score.size), max(int(cfg.n_best), 1),
                             alias_diverse=alias_diverse_for("BLS"))
    return period, freq, score, chosen, columns[chosen]


def _row_from_selection(prep: _Prepared, cand: _Candidates, cfg: _ChunkConfig, selection,
                        extras: np.ndarray) -> dict[str, Any]:
    period, freq, score, chosen, _ = selection
    return _row_of(prep, cfg, _peaks_from(period, freq, score, extras, chosen, cand.mean,
                                          cand.std))


def _candidate_row(prep: _Prepared, cand: _Candidates, settings: Any,
                   cfg: _ChunkConfig) -> dict[str, Any] | None:
    """The row :func:`periodogram_to_row` gives for this curve, from its device
    candidates. When they cannot decide the selection -- it would read past them, or
    equal scores among the visited ones could change it depending on how numpy sorts the
    whole candidate list -- the selection runs on the curve's full power row instead (one
    row, not seven), with the other outputs fetched at the chosen columns only; ``None``
    when even that is unavailable. SDE uses the device's mean and standard deviation of
    the power."""
    row = _row_from_candidates(prep, cand, settings, cfg)
    if row is not None:
        return row
    if cand.fetch_power is None or cand.fetch_extras is None or cand.periods_at is None:
        return None
    selection = _select_on_power(prep, cand, settings, cfg, cand.fetch_power())
    columns = selection[4]
    extras = cand.fetch_extras(columns) if columns.size else np.zeros((0, 4))
    return _row_from_selection(prep, cand, cfg, selection, extras)


def _candidate_rows(prepared: Sequence[_Prepared], cands: Sequence[_Candidates],
                    settings: Any, cfg: _ChunkConfig) -> list:
    """:func:`_candidate_row` for a launch's curves, with the power rows and the extras
    the undecided curves need read back in one transfer each. Per curve: the row,
    ``None`` (the full spectrum is needed) or the exception it raised."""
    out: list = [None] * len(prepared)
    undecided = []
    for i, (prep, cand) in enumerate(zip(prepared, cands)):
        if not isinstance(cand, _Candidates):
            continue
        try:
            out[i] = _row_from_candidates(prep, cand, settings, cfg)
        except Exception as exc:  # noqa: BLE001 - recorded per light curve
            out[i] = exc
            continue
        if out[i] is None and cand.reader is not None and cand.periods_at is not None:
            undecided.append(i)
    if not undecided:
        return out
    reader = cands[undecided[0]].reader
    powers = reader.power_rows([cands[i].curve for i in undecided])
    selections = {}
    for i in undecided:
        try:
            selections[i] = _select_on_power(prepared[i], cands[i], settings, cfg,
                                             powers[cands[i].curve])
        except Exception as exc:  # noqa: BLE001 - recorded per light curve
            out[i] = exc
    order = list(selections)
    extras = reader.extras([(cands[i].curve, selections[i][4]) for i in order])
    for i, ex in zip(order, extras):
        try:
            out[i] = _row_from_selection(prepared[i], cands[i], cfg, selections[i], ex)
        except Exception as exc:  # noqa: BLE001 - recorded per light curve
            out[i] = exc
    return out


class _ChunkRunner:
    """Packs and launches a chunk, then turns the previous launch into rows, so the
    host prepares chunk k+1 while the device searches chunk k."""

    def __init__(self, absorb: Callable[..., None]) -> None:
        self.absorb = absorb
        self.pending: tuple | None = None

    def submit(self, idx: int, items: Sequence[InputItem], cfg: _ChunkConfig) -> None:
        launches = []
        # position -> (index of the method that failed, key, message). As the
        # reference's chunk loop (src/cuperiod/batch/runner.py:220-249): a light curve's
        # methods run in order, a failure keeps the rows of the methods before it and
        # runs none after it.
        failed: dict[int, tuple[int, str, str]] = {}
        for mi, name in enumerate(cfg.methods):
            settings = cfg.settings_map[name]
            prepare = _prepare_tls if name == "TLS" else _prepare
            prepared: list[_Prepared] = []
            mb_rows: list[tuple[int, dict[str, Any]]] = []
            for pos, (key, lc) in enumerate(items):
                if pos in failed:
                    continue
                try:
                    if isinstance(lc, MultiBandLightCurve):
                        mb_rows.append((pos, _multiband_row(key, lc, name, settings, cfg)))
                    else:
                        prepared.append(prepare(pos, key, lc, settings, cfg.grid))
                except Exception as exc:  # noqa: BLE001 - one bad light curve
                    failed[pos] = (mi, key, f"{type(exc).__name__}: {exc}")
            if name == "TLS":
                handle = ("tls", _tls_rows(prepared, settings, cfg)) if prepared else None
            else:
                handle = _launch(prepared, settings, cfg.grid) if prepared else None
            launches.append((mi, settings, prepared, handle))
            if mb_rows:
                launches.append((mi, settings, [], ("rows", mb_rows)))
        self._complete()
        self.pending = (idx, cfg, launches, failed)

    def _complete(self) -> None:
        if self.pending is None:
            return
        idx, cfg, launches, failed = self.pending
        self.pending = None

        def runs(pos: int, mi: int) -> bool:
            # method mi ran on this light curve: no method before or at it failed
            return pos not in failed or mi < failed[pos][0]

        by_pos: dict[int, dict[int, dict[str, Any]]] = {}
        for mi, settings, prepared, handle in launches:
            if handle is None:
                continue
            if isinstance(handle[0], str):
                # rows computed at submission: ("tls", rows of `prepared`) or
                # ("rows", [(position, row)]) of multi-band light curves
                pairs = (zip((p.position for p in prepared), handle[1])
                         if handle[0] == "tls" else handle[1])
                for pos, row in pairs:
                    if runs(pos, mi):
                        by_pos.setdefault(pos, {})[mi] = row
                continue
            results = _collect(handle, candidates=not cfg.store_raw)
            from_candidates = (_candidate_rows(prepared, results, settings, cfg)
                               if not cfg.store_raw else [None] * len(prepared))
            for prep, res, row in zip(prepared, results, from_candidates):
                if not runs(prep.position, mi):
                    continue
                try:
                    if isinstance(row, Exception):
                        raise row
                    if row is None and isinstance(res, _Candidates):
                        res = res.fetch()
                    if row is None:
                        periods, spectrum = res
                        pg = _periodogram(prep, periods, spectrum, settings)
                        row = periodogram_to_row(prep.key, pg, n_best=cfg.n_best,
                                                 store_raw=cfg.store_raw,
                                                 extra_keys=cfg.extra_keys, source=prep.label)
                except Exception as exc:  # noqa: BLE001 - recorded per light curve
                    failed[prep.position] = (mi, prep.key, f"{type(exc).__name__}: {exc}")
                    continue
                by_pos.setdefault(prep.position, {})[mi] = row
        rows = [by_pos[pos][mi] for pos in sorted(by_pos) for mi in sorted(by_pos[pos])
                if runs(pos, mi)]
        errs = [(key, msg) for _, (_, key, msg) in sorted(failed.items())]
        self.absorb(idx, rows, errs)

    def drain(self) -> None:
        self._complete()


def names_of(method: str | Sequence[str]) -> tuple[str, ...]:
    """The requested method names, upper-cased."""
    names = (method,) if isinstance(method, str) else tuple(method)
    return tuple(str(n).upper() for n in names)


def _chunked(items: Sequence[Any], size: int) -> list[list[Any]]:
    return [list(items[i:i + size]) for i in range(0, len(items), size)]


def batch_periodograms(
    inputs: Any,
    method: str | Sequence[str] = "BLS",
    *,
    backend: str | None = "auto",
    device: str = "gpu",
    workers: int | None = None,
    grid: GridSpec | None = None,
    settings: Any = None,
    columns: Any = None,
    domain: Domain | str | None = None,
    band_column: str | None = None,
    sink: str | Path | None = None,
    n_best: int = 10,
    store_raw: bool = False,
    resume: bool = True,
    chunk_size: int | None = None,
) -> BatchSummary:
    """Periodograms of many light curves; one row per light curve and method.

    ``inputs`` is an iterable of :class:`LightCurve` objects or ``(key, lc)`` pairs.
    ``device="gpu"`` (the default here) runs the HIP engine in this process, one engine
    call per chunk of ``chunk_size`` light curves (default 256); ``device="cpu"``
    raises :class:`BackendUnavailableError`, since this package has no CPU backend.
    ``workers`` is accepted for signature compatibility: the batched engine fills the
    device from one process. ``grid`` may also be a grid rule, a callable that takes a
    light curve (its finite points) and returns its own grid -- a ``GridSpec`` or a
    parametric :class:`~cuperiod_amd.batch.engine.LinspaceGrid` (BLS only; not with a
    directory sink). ``grid``, ``settings``, ``sink``, ``n_best``,
    ``store_raw``, ``resume`` and ``chunk_size`` mean what they mean in cuPeriod
    (src/cuperiod/batch/runner.py:764-859). ``columns`` and ``domain`` only concern
    file and table inputs, which this package does not read; ``band_column``
    (multi-band input) is not supported.

    Raises
    ------
    ValueError
        Unknown device, a CPU backend with ``device="gpu"``, an unsupported sink, or a
        directory sink written by a different request.
    BackendUnavailableError
        When the HIP backend cannot run here (checked before any work).
    """
    if device not in {"cpu", "gpu", "hybrid"}:
        raise ValueError("device must be 'cpu', 'gpu', or 'hybrid'")
    if device == "hybrid":
        raise NotImplementedError("hybrid device is planned; use device='gpu'")
    if grid is not None and not isinstance(grid, GridSpec) and not callable(grid):
        raise TypeError(f"grid must be a GridSpec, a grid rule or None, got "
                        f"{type(grid).__name__}")
    if callable(grid) and not isinstance(grid, GridSpec):
        if set(names_of(method)) - {"BLS"}:
            raise ValueError("a grid rule (a callable grid) is supported for BLS only")
        if sink is not None and _classify_sink(sink)[0] == "dir":
            raise ValueError("a grid rule cannot be fingerprinted for a resumable directory "
                             "sink; pass a GridSpec")
    if chunk_size is not None and int(chunk_size) < 1:
        raise ValueError(f"chunk_size must be a positive integer, got {chunk_size}")
    if band_column is not None:
        raise NotImplementedError("multi-band batches are not supported by this package")
    domain = coerce_domain(domain)

    names = (method,) if isinstance(method, str) else tuple(method)
    resolved = [get_method(m) for m in names]
    method_objs = list({m.name: m for m in resolved}.values())
    if not method_objs:
        raise ValueError("no methods specified")
    methods = tuple(m.name for m in method_objs)
    settings_map = select_settings(method_objs, settings)
    backends = {n: _worker_backend(device, backend, settings_map[n]) for n in methods}
    for m in method_objs:
        m.resolve_backend_for(backends[m.name], settings_map[m.name])

    sink_kind, sink_dir, sink_file = _classify_sink(sink)
    if store_raw and sink_kind == "file" and sink_file.suffix.lower() == ".csv":
        raise ValueError(
            "store_raw=True produces array-valued spectrum columns that a CSV sink cannot "
            "hold; use a .parquet sink or a directory sink."
        )
    items = resolve_inputs(inputs)
    summary_sink = None if sink is None else str(sink)
    if not items:
        if summary_sink is not None:
            warnings.warn("batch_periodograms: nothing to process, so the sink was left "
                          "untouched (no inputs)", UserWarning, stacklevel=2)
        return BatchSummary(0, 0, 0, 0, methods, device, summary_sink,
                            rows=[] if sink_kind == "memory" else None, n_workers=0)

    multiband = any(isinstance(src, MultiBandLightCurve) for _, src in items)
    extra_keys = peak_extra_columns(methods, multiband=multiband)
    cfg = _ChunkConfig(methods, backends, settings_map, int(n_best), bool(store_raw),
                       grid, extra_keys)
    done: dict = {}
    next_index = 0
    csv_sink = sink_kind == "file" and sink_file.suffix.lower() == ".csv"
    if sink_kind == "dir":
        fingerprint = _run_fingerprint(
            methods, settings_map, n_best=n_best, store_raw=store_raw, columns=columns,
            domain=domain, band_column=band_column, grid=grid, extra_keys=extra_keys,
        )
        record = _dir_manifest_guard(sink_dir, resume, fingerprint)
        cfg = dataclasses.replace(cfg, extra_keys=tuple(record["config"]["extra_keys"]))
        if resume:
            done, next_index = _scan_parts(sink_dir)
    elif sink_kind == "file" and resume:
        done = _file_sink_done(sink_file)

    groups: dict[tuple[str, ...], list[InputItem]] = {}
    n_skipped = 0
    for key, src in items:
        lookup = _sink_key(key, csv_sink=csv_sink)
        missing = tuple(n for n in methods if (lookup, n) not in done)
        if missing:
            groups.setdefault(missing, []).append((key, src))
        else:
            n_skipped += 1

    size = int(chunk_size) if chunk_size is not None else DEFAULT_CHUNK_SIZE
    chunks: dict[int, list[InputItem]] = {}
    chunk_cfgs: dict[int, _ChunkConfig] = {}
    idx = next_index
    for missing, group in groups.items():
        group_cfg = cfg if missing == methods else dataclasses.replace(cfg, methods=missing)
        for chunk in _chunked(group, size):
            chunks[idx] = chunk
            chunk_cfgs[idx] = group_cfg
            idx += 1

    rows_by_chunk: dict[int, list[dict[str, Any]]] = {}
    errors_by_chunk: dict[int, list[tuple[str, str]]] = {}
    n_done = 0

    def absorb(i: int, rows: list[dict[str, Any]], errs: list[tuple[str, str]]) -> None:
        nonlocal n_done
        n_done += len(rows)
        errors_by_chunk[i] = errs
        if sink_kind == "dir":
            if rows:
                write_rows(rows, _part_path(sink_dir, i))
        else:
            rows_by_chunk[i] = rows

    runner = _ChunkRunner(absorb)
    for i in sorted(chunks):
        runner.submit(i, chunks[i], chunk_cfgs[i])
    runner.drain()

    def in_order(parts: Mapping[int, list]) -> list:
        return [entry for i in sorted(parts) for entry in parts[i]]

    kept_rows, errors = in_order(rows_by_chunk), in_order(errors_by_chunk)
    if sink_kind == "file":
        _finalize_file(kept_rows, sink_file, resume)
    return BatchSummary(len(items), n_done, len(errors), n_skipped, methods, device,
                        summary_sink, errors=errors,
                        rows=kept_rows if sink_kind == "memory" else None,
                        n_workers=int(bool(chunks)))


__all__ = ["DEFAULT_CHUNK_SIZE", "BatchSummary", "batch_periodograms"]
