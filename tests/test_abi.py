"""The C ABI: the library loads without a GPU and exports what the header declares."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from inundation_amd.hip import lib as L

HEADER = Path(__file__).resolve().parent.parent / "include" / "inundation_hip.h"


def _declared() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(inu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points() -> None:
    names = _declared()
    assert "inu_steps" in names and "inu_set_mesh" in names
    assert names == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol() -> None:
    lib = L.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing


def test_version_and_error_paths_need_no_gpu() -> None:
    lib = L.load()
    assert lib.inu_version() >= 1
    lib.inu_destroy(None)  # a null context is accepted
    assert lib.inu_steps(None, 1, 0) == L.EINVAL
    assert lib.inu_stream_copy(None, 16, 1, None) == L.EINVAL


def test_control_block_layout() -> None:
    assert C.sizeof(L.InuCtrl) == 48
    assert L.InuCtrl.dt_ticks.offset == 24 and L.InuCtrl.cfl_safety.offset == 40


@pytest.mark.parametrize("flag", ["FLAG_DETERMINISTIC", "FLAG_COMPENSATED", "FLAG_ACTIVE_SET", "FLAG_TRACK_CLAMP",
                                  "FLAG_EXACT_ORDER"])
def test_flags_are_distinct_bits(flag: str) -> None:
    v = getattr(L, flag)
    assert v and (v & (v - 1)) == 0
