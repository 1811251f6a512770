"""Known answers of the integer step clock, restated from the reference's kernel tests.

Each case is the reference's ``TestStepControl`` case
(``tests/unit/test_gpu_resident_clock.py:116-214``): a hand-built control
block (defaults of its harness, ``:72-77``: dt_max 1 s, dry dt 0.01 s, safety
0.8, stop at 10^9 ticks), the two dt words (the prologue bound and the
next-step bound, sentinel 1e30 by default) and the fields it checks after one
``step_control`` in the given mode. ``run_case(run, case)`` applies one to any
implementation: ``run(mode, ctrl, dt_words) -> (ctrl, dt_words)``.
"""

from __future__ import annotations

import struct
from collections.abc import Callable
from typing import Any

import numpy as np

from inundation_amd.hip.lib import InuCtrl
from inundation_amd.tuning import SIM_TICKS_PER_SECOND

START, STEP, RETARGET = 0, 1, 2
SENTINEL_BITS = struct.unpack("<I", struct.pack("<f", 1e30))[0]


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def shader_ticks(raw: float, safety: float = 0.8) -> int:
    """Ticks from a CFL bound in float32: trunc(f32(raw) * f32(safety) * 2^20) (reference :111-114)."""
    scaled = np.float32(raw) * np.float32(safety)
    return int(scaled * np.float32(SIM_TICKS_PER_SECOND))


def harness_ctrl(**fields: float) -> InuCtrl:
    c = InuCtrl()
    c.now_ticks = 0
    c.stop_ticks = 10**9
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = SIM_TICKS_PER_SECOND
    c.dry_dt_ticks = SIM_TICKS_PER_SECOND // 100
    c.last_raw = 0.0
    c.cfl_safety = 0.8
    for k, v in fields.items():
        setattr(c, k, v)
    return c


# (name, mode, slot0, slot1, control fields, expected fields; "dt0"/"dt1" are dt word bits)
CASES: list[tuple[str, int, float, float, dict[str, Any], dict[str, Any]]] = [
    ("start_takes_the_prologue_bound_with_safety", START, 0.05, 1e30, {"now_ticks": 123},
     {"dt_ticks": shader_ticks(0.05), "now_ticks": 123, "steps": 0,
      "dt0": f32_bits(shader_ticks(0.05) / SIM_TICKS_PER_SECOND), "dt1": SENTINEL_BITS}),
    ("step_advances_by_the_previous_ticks_and_counts", STEP, 0.0, 0.02,
     {"now_ticks": 1000, "dt_ticks": 777, "steps": 5},
     {"now_ticks": 1777, "steps": 6, "dt_ticks": shader_ticks(0.02), "last_raw": float(np.float32(0.02))}),
    ("noop_step_neither_advances_nor_counts_and_keeps_the_last_bound", STEP, 0.0, 1e30,
     {"now_ticks": 1000, "dt_ticks": 0, "steps": 5, "last_raw": 0.03},
     {"now_ticks": 1000, "steps": 5, "dt_ticks": shader_ticks(0.03)}),
    ("retarget_replans_from_the_last_bound", RETARGET, 0.0, 0.5,
     {"now_ticks": 1000, "dt_ticks": 0, "last_raw": 0.03},
     {"now_ticks": 1000, "dt_ticks": shader_ticks(0.03)}),
    ("capped_by_dt_max", START, 50.0, 1e30, {"dt_max_ticks": 4096}, {"dt_ticks": 4096}),
    ("last_step_lands_exactly_on_the_stop", START, 0.05, 1e30, {"now_ticks": 10_000, "stop_ticks": 10_005},
     {"dt_ticks": 5}),
    ("window_done_gives_a_zero_dt", STEP, 0.0, 0.05, {"now_ticks": 9_000, "dt_ticks": 1_000, "stop_ticks": 10_000},
     {"now_ticks": 10_000, "dt_ticks": 0, "dt0": 0}),
    ("all_dry_falls_back_to_the_dry_dt", START, 1e30, 1e30, {"dry_dt_ticks": 2048}, {"dt_ticks": 2048}),
    ("degenerate_bound_stalls_instead_of_stepping", START, 1e-12, 1e30, {}, {"dt_ticks": 0}),
    ("clock_carries_past_32_bits", STEP, 0.0, 0.05, {"now_ticks": 2**32 - 3, "dt_ticks": 10, "stop_ticks": 2**40},
     {"now_ticks": 2**32 - 3 + 10}),
    # Not among the reference's tests: its kernel caps every dt by dt_max after
    # choosing it (resident.py:877), so a dry fallback above dt_max is cut too.
    ("dry_fallback_is_capped_by_dt_max", START, 1e30, 1e30, {"dry_dt_ticks": 8192, "dt_max_ticks": 4096},
     {"dt_ticks": 4096}),
]
IDS = [c[0] for c in CASES]


def run_case(run: Callable[[int, InuCtrl, list[int]], tuple[InuCtrl, np.ndarray]], case: tuple) -> None:
    name, mode, slot0, slot1, fields, want = case
    ctrl, dt = run(mode, harness_ctrl(**fields), [f32_bits(slot0), f32_bits(slot1)])
    for key, value in want.items():
        got = int(dt[0]) if key == "dt0" else int(dt[1]) if key == "dt1" else getattr(ctrl, key)
        assert got == value, f"{name}: {key} = {got}, want {value}"


def dt_is_the_tick_count(run: Callable[[int, InuCtrl, list[int]], tuple[InuCtrl, np.ndarray]]) -> None:
    """Reference :131-137: the dt word times 2^20 is exactly the tick count."""
    ctrl, dt = run(START, harness_ctrl(), [f32_bits(0.0314), SENTINEL_BITS])
    dt_f = struct.unpack("<f", struct.pack("<I", int(dt[0])))[0]
    assert dt_f * SIM_TICKS_PER_SECOND == int(ctrl.dt_ticks)
    assert int(ctrl.dt_ticks) == shader_ticks(0.0314)
