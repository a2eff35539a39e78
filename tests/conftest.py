"""Test configuration: the ``gpu`` marker and import paths.

``-m gpu`` tests need a HIP device and the built gfx950 library; they are skipped at
setup when either is missing. Everything else runs on the CPU.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


def _hip_ready() -> bool:
    from cuperiod_amd import _lib

    return _lib.device_count() > 0


def pytest_runtest_setup(item: pytest.Item) -> None:
    if item.get_closest_marker("gpu") is not None and not _hip_ready():
        pytest.skip("no HIP device / library")
