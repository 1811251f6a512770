"""The integer clock on the GPU: k_step_control's known answers and whole runs through its branches.

``inu_probe_clock`` runs the device's step_control on a hand-set control
block, as the reference's own kernel tests do
(``tests/unit/test_gpu_resident_clock.py:116-214``, restated in
``tests/_clock_kats.py``). The runs drive a small mesh through the C ABI
into the branches the known answers pin (the dry fallback, the dt_max cap,
a clock past 2^32 ticks, exact landing on the stop) and compare the device
with the FP32 oracle bit for bit, step by step.
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays
from inundation_amd.tuning import dt_to_ticks
from oracle import cpu_ref
from tests import _clock_kats as K
from tests import _drive as D
from tests._meshes import MANNING, f32_state, valley_grid, wet_lake
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():  # noqa: ANN201
    c = gpu_context()
    yield c
    c.close()


@pytest.mark.parametrize("case", K.CASES, ids=K.IDS)
def test_device_step_control_known_answers(ctx, case: tuple) -> None:  # noqa: ANN001
    print(f"clock case {case[0]}", flush=True)
    K.run_case(ctx.probe_clock, case)


def test_device_dt_word_is_exactly_the_tick_count(ctx) -> None:  # noqa: ANN001
    K.dt_is_the_tick_count(ctx.probe_clock)


def test_device_and_oracle_agree_on_random_control_blocks(ctx) -> None:  # noqa: ANN001
    rng = np.random.default_rng(7)
    for k in range(300):
        mode = int(rng.integers(0, 3))
        now = int(rng.choice([0, 2**32 - int(rng.integers(1, 5000)), int(rng.integers(0, 2**40))]))
        fields = {
            "now_ticks": now,
            "stop_ticks": now + int(rng.choice([0, int(rng.integers(1, 2**22)), 2**40])),
            "steps": int(rng.integers(0, 1000)),
            "dt_ticks": int(rng.choice([0, int(rng.integers(1, 2**20))])),
            "dt_max_ticks": int(rng.integers(1, 2**24)),
            "dry_dt_ticks": int(rng.integers(1, 2**24)),
            "last_raw": float(10.0 ** rng.uniform(-13, 31)),
            "cfl_safety": float(rng.uniform(0.1, 1.0)),
        }
        slots = [K.f32_bits(float(10.0 ** rng.uniform(-13, 31))) for _ in range(2)]
        a_ctrl, a_dt = ctx.probe_clock(mode, K.harness_ctrl(**fields), slots)
        b_ctrl, b_dt = cpu_ref.probe_clock(mode, K.harness_ctrl(**fields), slots)
        assert bytes(a_ctrl) == bytes(b_ctrl), (k, mode, fields)
        assert np.array_equal(a_dt, b_dt), (k, mode, fields)


@pytest.fixture(scope="module")
def grid():  # noqa: ANN201
    return valley_grid(64, 32, 2.0)


def _pair(ctx, grid):  # noqa: ANN001, ANN202
    m = device_mesh_arrays(grid)
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    return ctx, cpu_ref.CpuRef(m)


def _ctrl(now: int, stop: int, dt_max: float, dry_dt: float) -> L.InuCtrl:
    c = L.InuCtrl()
    c.now_ticks = now
    c.stop_ticks = stop
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(dt_max)
    c.dry_dt_ticks = dt_to_ticks(dry_dt)
    c.cfl_safety = 1.0
    return c


def _run_both(ctx, ref, grid, state, ctrl, steps: int) -> list[tuple[int, int]]:  # noqa: ANN001
    nb = -(-grid.N // 256)
    for be in (ctx, ref):
        D.setup(be, grid.N, *state, MANNING, dry_tol=1e-3)
        be.ctrl_write(ctrl)
        be.prologue()
    clock = []
    for k in range(steps):
        for be in (ctx, ref):
            be.steps(1) if be is ref else be.steps(1, use_graph=False)
        got, want = D.snapshot(ctx, grid.N, nb), D.snapshot(ref, grid.N, nb)
        D.assert_same(got, want, f"step {k + 1}")
        c = ctx.ctrl_read()
        clock.append((int(c.now_ticks), int(c.dt_ticks)))
    return clock


def test_an_all_dry_mesh_steps_at_the_dry_dt(ctx, grid) -> None:  # noqa: ANN001
    gpu, ref = _pair(ctx, grid)
    z = np.zeros(grid.N, np.float32)
    clock = _run_both(gpu, ref, grid, (z, z, z), _ctrl(0, 2**40, 1.0, 0.25), 4)
    assert [dt for _, dt in clock] == [dt_to_ticks(0.25)] * 4
    assert clock[-1][0] == 4 * dt_to_ticks(0.25)


def test_the_dt_max_cap_holds_and_the_clock_carries_past_32_bits(ctx, grid) -> None:  # noqa: ANN001
    gpu, ref = _pair(ctx, grid)
    state = f32_state(*wet_lake(grid))
    dt_max = 2.0**-12  # far below the CFL bound of this lake
    now = 2**32 - 3 * dt_to_ticks(dt_max) - 5
    clock = _run_both(gpu, ref, grid, state, _ctrl(now, 2**40, dt_max, 0.01), 6)
    assert all(dt == dt_to_ticks(dt_max) for _, dt in clock)
    assert clock[-1][0] == now + 6 * dt_to_ticks(dt_max) > 2**32


def test_the_last_step_lands_on_the_stop(ctx, grid) -> None:  # noqa: ANN001
    gpu, ref = _pair(ctx, grid)
    state = f32_state(*wet_lake(grid))
    stop = 3 * dt_to_ticks(2.0**-12) + 7
    clock = _run_both(gpu, ref, grid, state, _ctrl(0, stop, 2.0**-12, 0.01), 5)
    assert clock[2][0] == stop - 7 and clock[3][0] == stop and clock[4] == (stop, 0)
