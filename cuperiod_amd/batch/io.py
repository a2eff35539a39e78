"""Batch inputs, result rows and sinks.

The row format and sink behaviour of cuPeriod's batch driver
(src/cuperiod/batch/io.py:58-131, 395-482, 507-667): one row per light curve and
method with ``key, method, source, backend, n_samples, baseline, best_period`` and, for
ranks ``k = 1..n_best``, ``period_k, power_k`` and the method's declared per-peak extras
``{extra}_k`` (NaN past the last peak); BLS peaks are chosen alias-diverse. Parquet and
CSV files are written to a new hidden file beside the target and renamed over it; CSV
text cells that a spreadsheet would evaluate are quote-prefixed.
"""

from __future__ import annotations

import csv
import numbers
import os
import secrets
import warnings
from collections import Counter
from collections.abc import Iterable, Mapping, Sequence
from pathlib import Path
from typing import Any

from cuperiod_amd.core.lightcurve import LightCurve, MultiBandLightCurve
from cuperiod_amd.core.result import Peak, Periodogram

#: Methods whose secondary peaks skip harmonics and aliases of the chosen ones.
ALIAS_DIVERSE_METHODS: frozenset[str] = frozenset({"BLS", "TLS"})

#: Per-peak extras each method declares, in column order.
PEAK_EXTRA_COLUMNS: Mapping[str, tuple[str, ...]] = {
    "BLS": ("depth", "depth_snr", "duration", "t0", "sde"),
    "TLS": ("sr", "depth", "duration", "t0"),
}

#: Extras only multi-band results carry (src/cuperiod/batch/io.py:73-75).
MULTIBAND_PEAK_EXTRA_COLUMNS: Mapping[str, tuple[str, ...]] = {
    "BLS": ("band_index",),
}

_CSV_FORMULA_PREFIXES = ("=", "+", "-", "@", "\t", "\r")

InputItem = tuple[str, Any]


def _canonical_method(name: str) -> str:
    return "".join(ch for ch in str(name) if ch.isalnum()).upper()


def alias_diverse_for(method: str) -> bool:
    """Whether ``method``'s batch peaks are selected alias-diverse."""
    return _canonical_method(method) in ALIAS_DIVERSE_METHODS


def peak_extra_columns(methods: Sequence[str], *, multiband: bool = False) -> tuple[str, ...]:
    """Ordered union of the per-peak extras ``methods`` declare; ``multiband`` adds the
    extras only multi-band results carry."""
    tables = (PEAK_EXTRA_COLUMNS, MULTIBAND_PEAK_EXTRA_COLUMNS) if multiband \
        else (PEAK_EXTRA_COLUMNS,)
    names = (extra for method in map(_canonical_method, methods)
             for table in tables for extra in table.get(method, ()))
    return tuple(dict.fromkeys(names))


def source_label(source: Any) -> str:
    """Provenance of an input: an absolute path, or ``""`` for an in-memory curve."""
    if isinstance(source, (str, os.PathLike)):
        return str(Path(source).resolve())
    return ""


def _is_key(value: object) -> bool:
    return isinstance(value, str) or (
        isinstance(value, numbers.Integral) and not isinstance(value, bool)
    )


def resolve_inputs(inputs: Any) -> list[InputItem]:
    """``[(key, source)]`` from an iterable of light curves (single- or multi-band) or
    ``(key, lc)`` pairs.

    A bare light curve is keyed by ``meta["id"]`` or its position. Keys must be strings
    or integers; repeated keys are renamed ``key#2``, ``key#3``, ... with a warning.
    File, glob and table inputs are not read by this package: load them into
    :class:`LightCurve` / :class:`MultiBandLightCurve` objects first.
    """
    if isinstance(inputs, (str, os.PathLike)):
        raise TypeError(
            "batch_periodograms: file and glob inputs are not read by this package; pass "
            "LightCurve objects or (key, LightCurve) pairs"
        )
    sources = (LightCurve, MultiBandLightCurve)
    if isinstance(inputs, sources):
        inputs = [inputs]
    items: list[InputItem] = []
    for i, entry in enumerate(inputs):
        if isinstance(entry, tuple) and len(entry) == 2 and isinstance(entry[1], sources):
            key, lc = entry
            if not _is_key(key):
                raise TypeError(
                    f"batch key at index {i} must be a str or an integer, got "
                    f"{type(key).__name__}"
                )
            items.append((str(key), lc))
        elif isinstance(entry, sources):
            key = str(entry.meta.get("id", i)) if entry.meta else str(i)
            items.append((key, entry))
        elif isinstance(entry, tuple) and len(entry) == 2 and isinstance(
            entry[1], (str, os.PathLike)
        ):
            raise TypeError(
                f"batch entry at index {i} names a file; this package reads light curves "
                "from memory only"
            )
        else:
            raise TypeError(f"unsupported batch entry at index {i}: {type(entry)!r}")
    return _unique_keys(items)


def _unique_keys(items: list[InputItem]) -> list[InputItem]:
    """Repeated keys (compared as a CSV sink stores them) renamed ``key#n``."""
    stored = [str(csv_safe(key)) for key, _ in items]
    if all(n == 1 for n in Counter(stored).values()):
        return items
    taken = set(stored)
    seen: Counter[str] = Counter()
    out: list[InputItem] = []
    renamed: list[str] = []
    for (key, src), form in zip(items, stored):
        seen[form] += 1
        if seen[form] == 1:
            out.append((key, src))
            continue
        n = seen[form]
        while str(csv_safe(f"{key}#{n}")) in taken:
            n += 1
        new = f"{key}#{n}"
        taken.add(str(csv_safe(new)))
        out.append((new, src))
        renamed.append(new)
    warnings.warn(
        f"{len(renamed)} batch keys occur more than once; the repeats were renamed "
        f"({', '.join(repr(k) for k in renamed[:3])}) so each light curve keeps its own "
        "result rows and resume entry.",
        UserWarning,
        stacklevel=3,
    )
    return out


def periodogram_to_row(
    key: str,
    pg: Periodogram,
    *,
    n_best: int,
    store_raw: bool,
    extra_keys: Sequence[str] | None = None,
    alias_diverse: bool | None = None,
    source: str | None = None,
) -> dict[str, Any]:
    """One result row: identity columns, ``best_period``, then per-rank peaks."""
    diverse = alias_diverse_for(pg.method) if alias_diverse is None else alias_diverse
    peaks = pg.best_periods(max(int(n_best), 1), alias_diverse=diverse)
    raw = pg.downsample() if store_raw else None
    return peaks_to_row(key, pg.method, pg.backend, pg.n_samples, pg.baseline, peaks,
                        n_best=n_best, extra_keys=extra_keys, source=source, raw=raw)


def peaks_to_row(
    key: str,
    method: str,
    backend: str,
    n_samples: int,
    baseline: float,
    peaks: Sequence[Peak],
    *,
    n_best: int,
    extra_keys: Sequence[str] | None = None,
    source: str | None = None,
    raw: tuple | None = None,
) -> dict[str, Any]:
    """The row of :func:`periodogram_to_row` from already selected peaks (most
    significant first) and, for ``store_raw``, the downsampled ``(frequency, power)``."""
    nan = float("nan")
    declared = () if extra_keys is None else tuple(extra_keys)
    row: dict[str, Any] = {"key": key, "method": method}
    row.update({} if source is None else {"source": source})
    row.update(backend=backend, n_samples=n_samples, baseline=baseline,
               best_period=peaks[0].period if peaks else nan)
    for rank in range(1, n_best + 1):
        peak = peaks[rank - 1] if rank <= len(peaks) else None
        extra = {} if peak is None else peak.extra
        row[f"period_{rank}"] = nan if peak is None else peak.period
        row[f"power_{rank}"] = nan if peak is None else peak.power
        # declared extras always get a (float) column; undeclared ones as they come
        row.update({f"{name}_{rank}": nan if extra.get(name) is None else float(extra[name])
                    for name in declared})
        row.update({f"{name}_{rank}": value for name, value in extra.items()
                    if name not in declared})
    if raw is not None:
        row["pgram_frequency"], row["pgram_power"] = (a.tolist() for a in raw)
    return row


def _numeric_text(text: str) -> bool:
    try:
        float(text)
    except ValueError:
        return False
    return True


def csv_safe(value: Any) -> Any:
    """Quote-prefix a text cell starting with ``= + - @ TAB CR`` unless it is a number."""
    risky = isinstance(value, str) and value[:1] in _CSV_FORMULA_PREFIXES and value != ""
    return "'" + value if risky and not _numeric_text(value) else value


def _union_columns(rows: Sequence[Mapping[str, Any]]) -> list[str]:
    seen: dict[str, None] = {}
    for row in rows:
        for name in row:
            seen.setdefault(name, None)
    return list(seen)


def _publish(path: Path, mode: str, fill) -> None:
    """Write a file through ``fill(handle)`` into a new hidden sibling of ``path``, then
    rename it over ``path``: a reader sees the old file or the complete new one. A failed
    write leaves nothing behind; a failed rename leaves the complete file where the
    error message says."""
    staging = path.parent / f".{path.name}.{os.getpid()}-{secrets.token_hex(4)}.part"
    kwargs = {} if "b" in mode else {"newline": "", "encoding": "utf-8"}
    try:
        with open(staging, mode, **kwargs) as handle:
            fill(handle)
    except BaseException:
        staging.unlink(missing_ok=True)
        raise
    try:
        os.replace(staging, path)
    except OSError as exc:
        raise OSError(exc.errno, f"result file {path} not replaced ({exc.strerror}); "
                      f"the finished rows are in {staging}") from exc


def _write_parquet(rows: Sequence[Mapping[str, Any]], path: Path) -> None:
    import pyarrow as pa
    import pyarrow.parquet as pq

    names = _union_columns(rows)
    table = pa.Table.from_pylist([{c: r.get(c) for c in names} for r in rows])
    _publish(path, "xb", lambda fh: pq.write_table(table, fh))


def _write_csv(rows: Sequence[Mapping[str, Any]], path: Path) -> None:
    names = _union_columns(rows)
    arrays = sorted({c for r in rows for c, v in r.items() if isinstance(v, list)})
    if arrays:
        raise ValueError(f"a CSV sink holds scalars only, and the rows carry arrays in "
                         f"{arrays} (store_raw): write to a .parquet sink")

    def fill(fh) -> None:
        out = csv.writer(fh)
        out.writerow(names)
        for r in rows:
            out.writerow([csv_safe(r.get(c, "")) for c in names])

    _publish(path, "x", fill)


def write_rows(rows: Sequence[Mapping[str, Any]], path: Path) -> None:
    """Write rows to ``path`` (``.parquet``/``.pq`` or ``.csv``) atomically."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    suffix = path.suffix.lower()
    if suffix in {".parquet", ".pq"}:
        _write_parquet(rows, path)
    elif suffix == ".csv":
        _write_csv(rows, path)
    else:
        raise ValueError(f"unsupported sink format {path.suffix!r}; use .parquet or .csv")


def read_rows(path: Path) -> list[dict[str, Any]]:
    """All rows of a ``.parquet`` or ``.csv`` file."""
    path = Path(path)
    if path.suffix.lower() in {".parquet", ".pq"}:
        import pyarrow.parquet as pq

        with path.open("rb") as fh:
            return pq.read_table(fh).to_pylist()
    with path.open(newline="", encoding="utf-8") as fh:
        return [dict(row) for row in csv.DictReader(fh)]


__all__ = [
    "ALIAS_DIVERSE_METHODS",
    "MULTIBAND_PEAK_EXTRA_COLUMNS",
    "PEAK_EXTRA_COLUMNS",
    "InputItem",
    "alias_diverse_for",
    "csv_safe",
    "peak_extra_columns",
    "peaks_to_row",
    "periodogram_to_row",
    "read_rows",
    "resolve_inputs",
    "source_label",
    "write_rows",
]
