"""Test configuration: the ``gpu`` marker and the HIP availability gate."""

from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


def gpu_context():
    """A HIP context, or fail loudly: a GPU test must never pass on a fallback."""
    from inundation_amd.hip import lib

    return lib.Context(int(os.environ.get("INU_DEVICE", "0")))
