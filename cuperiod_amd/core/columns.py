"""Brightness domain of a light curve (src/cuperiod/core/columns.py:34-67)."""

from __future__ import annotations

import enum


class Domain(str, enum.Enum):
    """Magnitude or flux. Built from a string ignoring case and surrounding blanks."""

    MAGNITUDE = "magnitude"
    FLUX = "flux"

    @classmethod
    def _missing_(cls, value: object) -> "Domain | None":
        if isinstance(value, str):
            norm = value.strip().lower()
            for member in cls:
                if member.value == norm:
                    return member
        return None

    def __str__(self) -> str:  # StrEnum behaviour on Python 3.10
        return self.value


def coerce_domain(domain: "Domain | str | None") -> "Domain | None":
    """``None`` stays ``None``; a string or :class:`Domain` becomes a :class:`Domain`."""
    return None if domain is None else Domain(domain)


__all__ = ["Domain", "coerce_domain"]
