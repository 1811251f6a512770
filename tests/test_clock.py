"""The integer clock on the CPU: the oracle's step_control and the host tick helpers.

The step_control cases are the reference's known answers
(``tests/unit/test_gpu_resident_clock.py:116-214``, restated in
``tests/_clock_kats.py``); the tick cases restate
``tests/unit/test_sim_ticks.py:28-114`` against ``inundation_amd.tuning``,
including the exhaustive float32 check of every tick count up to 2^24. A
last group plants plausible mistakes in a Python model of step_control and
shows that each one fails at least one known answer, so the cases pin the
branches they are meant to pin.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from inundation_amd.tuning import (
    SIM_DT_MAX_TICKS,
    SIM_TICKS_PER_SECOND,
    dt_to_ticks,
    output_window_stops,
    seconds_to_ticks,
    ticks_to_seconds,
)
from oracle import cpu_ref
from tests import _clock_kats as K

DAY_S = 86_400.0


# -- the oracle's step_control ------------------------------------------------


@pytest.mark.parametrize("case", K.CASES, ids=K.IDS)
def test_oracle_step_control_known_answers(case: tuple) -> None:
    K.run_case(cpu_ref.probe_clock, case)


def test_oracle_dt_word_is_exactly_the_tick_count() -> None:
    K.dt_is_the_tick_count(cpu_ref.probe_clock)


# -- planted mistakes: each fails a known answer -------------------------------


def _model(mistake: str | None):  # noqa: ANN202
    """step_control in Python (float32 where the kernel is), with one optional planted mistake."""
    f32 = np.float32

    def run(mode: int, ctrl, dt_words):  # noqa: ANN001, ANN202
        c = type(ctrl).from_buffer_copy(ctrl)
        slot0, slot1 = (np.array(dt_words[:2], dtype=np.uint32).view(np.float32)).tolist()
        if mode == K.START:
            raw = f32(slot0)
            c.last_raw = raw
        elif mode == K.STEP:
            done = int(c.dt_ticks)
            c.now_ticks = (int(c.now_ticks) + done) % (2**64 if mistake != "clock_wraps_at_32_bits" else 2**32)
            if done > 0 or mistake == "noop_counts":
                c.steps += 1
                raw = f32(slot1)
                c.last_raw = raw
            else:
                raw = f32(c.last_raw)
        else:
            raw = f32(c.last_raw) if mistake != "retarget_reads_next_slot" else f32(slot1)
        remaining = max(int(c.stop_ticks) - int(c.now_ticks), 0)
        if remaining == 0 or raw < f32(1e-10):
            ticks = 0 if mistake != "no_stall" else int(c.dry_dt_ticks)
        elif raw >= f32(1e10):
            ticks = int(c.dry_dt_ticks) if mistake != "dry_takes_dt_max" else int(c.dt_max_ticks)
        else:
            safety = f32(c.cfl_safety) if mistake != "no_safety" else f32(1.0)
            cap = f32(c.dt_max_ticks) / f32(SIM_TICKS_PER_SECOND)
            scaled = min(raw * safety, cap) * f32(SIM_TICKS_PER_SECOND)
            ticks = int(np.rint(scaled)) if mistake == "round_not_floor" else int(scaled)
        if mistake != "no_cap":
            ticks = min(ticks, int(c.dt_max_ticks))
        if ticks > remaining and mistake != "overshoots_stop":
            ticks = remaining
        c.dt_ticks = ticks
        dt = np.array([f32(ticks) / f32(SIM_TICKS_PER_SECOND), f32(1e30), 0, 0], dtype=np.float32).view(np.uint32)
        return c, dt

    return run


def test_the_python_model_passes_every_known_answer() -> None:
    for case in K.CASES:
        K.run_case(_model(None), case)


@pytest.mark.parametrize(
    "mistake",
    ["dry_takes_dt_max", "no_safety", "round_not_floor", "no_stall", "no_cap", "overshoots_stop",
     "clock_wraps_at_32_bits", "noop_counts", "retarget_reads_next_slot"],
)
def test_a_planted_mistake_fails_a_known_answer(mistake: str) -> None:
    failed = []
    for case in K.CASES:
        try:
            K.run_case(_model(mistake), case)
        except AssertionError:
            failed.append(case[0])
    assert failed, f"no known answer catches {mistake}"


# -- tick helpers (reference tests/unit/test_sim_ticks.py) ----------------------


@pytest.mark.parametrize("seconds", [0.0, 0.5, 1.0, 600.0, 1800.0, 7200.0, DAY_S])
def test_phase_and_output_times_are_exact(seconds: float) -> None:
    ticks = seconds_to_ticks(seconds)
    assert ticks == int(seconds * SIM_TICKS_PER_SECOND)
    assert ticks_to_seconds(ticks) == seconds


def test_seconds_round_to_the_nearest_tick() -> None:
    ticks = seconds_to_ticks(0.1)
    assert ticks == round(0.1 * SIM_TICKS_PER_SECOND)
    assert abs(ticks_to_seconds(ticks) - 0.1) <= 0.5 / SIM_TICKS_PER_SECOND


def test_summed_phases_land_on_the_total() -> None:
    durations = [600.0] * 12 + [79_200.0]
    assert sum(seconds_to_ticks(d) for d in durations) == seconds_to_ticks(DAY_S)


@pytest.mark.parametrize("bad", [-1.0, math.nan, math.inf, -math.inf])
def test_seconds_reject_negative_and_non_finite(bad: float) -> None:
    with pytest.raises(ValueError, match="simulated time"):
        seconds_to_ticks(bad)


def test_a_year_fits_in_uint64() -> None:
    assert seconds_to_ticks(365 * DAY_S) < 2**64


def test_dt_rounds_down_so_the_cfl_bound_is_never_exceeded() -> None:
    dt = 0.0314
    ticks = dt_to_ticks(dt)
    assert ticks_to_seconds(ticks) <= dt
    assert dt - ticks_to_seconds(ticks) < 1.0 / SIM_TICKS_PER_SECOND


def test_dt_upper_limit_is_sixteen_seconds() -> None:
    assert dt_to_ticks(16.0) == SIM_DT_MAX_TICKS
    with pytest.raises(ValueError, match="dt"):
        dt_to_ticks(16.5)


@pytest.mark.parametrize("bad", [0.0, -0.1, math.nan, math.inf])
def test_dt_rejects_non_positive_and_non_finite(bad: float) -> None:
    with pytest.raises(ValueError, match="dt"):
        dt_to_ticks(bad)


def test_dt_rejects_below_one_tick() -> None:
    with pytest.raises(ValueError, match="dt"):
        dt_to_ticks(0.5 / SIM_TICKS_PER_SECOND)


def test_every_allowed_tick_count_is_exact_in_float32() -> None:
    # The kernel forms dt as float(ticks) / 2^20 in float32: exhaustive over
    # [1, 2^24], both the integer and the scaling must be exact; also as a
    # product with 2^-20, the reference's form.
    ticks = np.arange(1, SIM_DT_MAX_TICKS + 1, dtype=np.int64)
    exact = ticks / SIM_TICKS_PER_SECOND
    as_f32 = ticks.astype(np.float32) * np.float32(1.0 / SIM_TICKS_PER_SECOND)
    assert np.array_equal(as_f32.astype(np.float64), exact)
    divided = ticks.astype(np.float32) / np.float32(SIM_TICKS_PER_SECOND)
    assert np.array_equal(divided.astype(np.float64), exact)


def test_outputs_every_interval_then_the_end() -> None:
    stops = list(output_window_stops(0, 3600 * SIM_TICKS_PER_SECOND, 1800 * SIM_TICKS_PER_SECOND))
    assert stops == [1800 * SIM_TICKS_PER_SECOND, 3600 * SIM_TICKS_PER_SECOND]


def test_last_window_is_capped_at_the_end() -> None:
    assert list(output_window_stops(0, 1000, 600)) == [600, 1000]


def test_interval_longer_than_the_run_gives_one_window() -> None:
    assert list(output_window_stops(0, 500, 600)) == [500]


def test_windows_are_relative_to_the_run_start() -> None:
    start = seconds_to_ticks(7200.0)
    stops = list(output_window_stops(start, start + seconds_to_ticks(3600.0), seconds_to_ticks(1800.0)))
    assert stops == [start + seconds_to_ticks(1800.0), start + seconds_to_ticks(3600.0)]


def test_no_window_when_already_at_the_end() -> None:
    assert list(output_window_stops(1000, 1000, 600)) == []


def test_windows_reject_a_non_positive_interval() -> None:
    with pytest.raises(ValueError, match="interval"):
        list(output_window_stops(0, 1000, 0))
