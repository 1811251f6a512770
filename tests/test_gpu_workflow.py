"""The workflow's output layer on the HIP backend (needs an MI355X).

Reference ``tests/unit/test_gpu_momentum_snapshots.py``: ``hu_snapshots`` and
``hv_snapshots`` line up one for one with the depth snapshots across phases
(a phase boundary's snapshot appears twice, identically), the first is the
initial state and the last the final state; capturing them changes no bit of
the run; a snapshot equals the state of a run stopped at that time; blocks
that dried and went inactive hold zero momentum, as with the active set off.
Plus warm starts: a run saved with ``save_warm_start`` and resumed from the
file continues bit for bit as the uninterrupted run.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from inundation_amd.hip.solver import SWESolverHipResident
from inundation_amd.tuning import ACTIVE_BLOCK_CELLS
from inundation_amd.workflow import (
    PointSource,
    SimulationPhase,
    SWEWorkflow,
    WorkflowConfig,
    WorkflowResult,
    save_warm_start,
)
from tests._meshes import MANNING, sheet, valley_grid

pytestmark = pytest.mark.gpu

OUTPUT_S = 5.0
PHASES_S = (12.0, 8.0)  # both end between outputs
LAST_ON_OUTPUT_S = 10.0
LAST_BETWEEN_S = 13.0
SOURCE_XY = (10.0, 48.0)


def _sloped_obj(path: Path, n: int = 48, dx: float = 2.0) -> None:
    rows = [f"v {i * dx} {j * dx} {10.0 - 0.02 * i * dx}" for j in range(n + 1) for i in range(n + 1)]

    def vid(i: int, j: int) -> int:
        return j * (n + 1) + i + 1

    rows += [f"f {vid(i, j)} {vid(i + 1, j)} {vid(i + 1, j + 1)} {vid(i, j + 1)}" for j in range(n) for i in range(n)]
    path.write_text("\n".join(rows) + "\n", encoding="ascii")


@pytest.fixture(scope="module")
def mesh(tmp_path_factory) -> Path:  # noqa: ANN001
    p = tmp_path_factory.mktemp("workflow") / "sloped.obj"
    _sloped_obj(p)
    return p


def _workflow(mesh: Path, capture: bool, **cfg) -> SWEWorkflow:  # noqa: ANN003
    wf = SWEWorkflow(
        WorkflowConfig(
            mesh_source=str(mesh),
            manning_n=MANNING,
            output_interval_s=OUTPUT_S,
            dt_max=1.0,
            cfl_interval=10,
            progress=False,
            capture_momentum_snapshots=capture,
            **cfg,
        )
    )
    wf.prepare()
    return wf


def _phases(last_s: float) -> list[SimulationPhase]:
    return [
        SimulationPhase(PHASES_S[0], [PointSource(2.0, SOURCE_XY, 3.0)]),
        SimulationPhase(PHASES_S[1], [PointSource(6.0, SOURCE_XY, 3.0)]),
        SimulationPhase(last_s, []),
    ]


def _run(mesh: Path, capture: bool, last_s: float) -> WorkflowResult:
    wf = _workflow(mesh, capture)
    res = wf.run(_phases(last_s))
    wf.solver.close()
    return res


@pytest.fixture(scope="module")
def captured(mesh: Path) -> WorkflowResult:
    return _run(mesh, True, LAST_ON_OUTPUT_S)


@pytest.fixture(scope="module")
def captured_tail(mesh: Path) -> WorkflowResult:
    return _run(mesh, True, LAST_BETWEEN_S)


@pytest.mark.parametrize("which", ["captured", "captured_tail"])
def test_momentum_lists_align_across_phases(which: str, request) -> None:  # noqa: ANN001
    r: WorkflowResult = request.getfixturevalue(which)
    assert r.hu_snapshots is not None and r.hv_snapshots is not None
    assert len(r.hu_snapshots) == len(r.hv_snapshots) == len(r.snapshots) == len(r.snap_times)
    assert max(float(np.abs(a).max()) for a in r.hu_snapshots) > 0.0
    assert max(float(np.abs(a).max()) for a in r.hv_snapshots) > 0.0


def test_a_phase_boundary_snapshot_appears_twice(captured: WorkflowResult) -> None:
    t = np.asarray(captured.snap_times)
    for boundary in (12.0, 20.0):
        first, second = (int(k) for k in np.flatnonzero(t == boundary))
        assert second == first + 1
        for series in (captured.snapshots, captured.hu_snapshots, captured.hv_snapshots):
            assert np.array_equal(series[first], series[second])


def test_first_snapshot_is_the_initial_state(captured: WorkflowResult) -> None:
    assert captured.snap_times[0] == 0.0
    assert not captured.hu_snapshots[0].any() and not captured.hv_snapshots[0].any()


@pytest.mark.parametrize(("which", "last"), [("captured", LAST_ON_OUTPUT_S), ("captured_tail", LAST_BETWEEN_S)])
def test_last_snapshot_is_the_final_state(which: str, last: float, request) -> None:  # noqa: ANN001
    r: WorkflowResult = request.getfixturevalue(which)
    assert r.snap_times[-1] == sum(PHASES_S) + last
    assert np.array_equal(r.hu_snapshots[-1], r.hu_final)
    assert np.array_equal(r.hv_snapshots[-1], r.hv_final)
    assert np.array_equal(r.snapshots[-1], r.h_final)


def test_capturing_momentum_changes_no_bit(mesh: Path, captured_tail: WorkflowResult) -> None:
    off = _run(mesh, False, LAST_BETWEEN_S)
    assert off.hu_snapshots is None and off.hv_snapshots is None
    assert off.snap_times == captured_tail.snap_times
    assert off.steps_total == captured_tail.steps_total
    for a, b in zip([*off.snapshots, off.h_final], [*captured_tail.snapshots, captured_tail.h_final], strict=True):
        assert a.tobytes() == b.tobytes()


def test_point_source_volume_arrives(captured: WorkflowResult) -> None:
    injected = 2.0 * PHASES_S[0] + 6.0 * PHASES_S[1]
    assert captured.volume_injected_m3 == pytest.approx(injected)
    assert captured.volume_final_m3 == pytest.approx(injected, rel=1e-5)


# ---- the solver's state at output times ------------------------------------


@pytest.fixture(scope="module")
def grid():  # noqa: ANN201
    return valley_grid(256, 128, 2.0)


def _solver(grid, h0, **kw) -> SWESolverHipResident:  # noqa: ANN001, ANN003
    z = np.zeros(grid.N, dtype=np.float32)
    return SWESolverHipResident(grid, None, h0, z, z, np.full(grid.N, MANNING, np.float32), **kw)


def _near(grid, x: float, y: float) -> int:  # noqa: ANN001
    return int(np.argmin((grid.centroid[:, 0] - x) ** 2 + (grid.centroid[:, 1] - y) ** 2))


def _go(s: SWESolverHipResident, t_end: float, interval: float, **kw):  # noqa: ANN003, ANN202
    return s.run(t_end=t_end, output_interval_s=interval, dt_max=1.0, dt_init=0.01, cfl_interval=10,
                 progress=False, **kw)


def test_a_snapshot_is_the_state_of_a_run_stopped_then(grid) -> None:  # noqa: ANN001
    src = np.zeros(grid.N, dtype=np.float32)
    src[_near(grid, 256 * 2.0 * 0.2, 128 * 2.0 / 2)] = 1.0
    h0 = np.zeros(grid.N, dtype=np.float32)
    s = _solver(grid, h0)
    _, times, hu, hv = _go(s, 30.0, 10.0, source_rate=src, capture_momentum_snapshots=True)
    s.close()
    assert times == [0.0, 10.0, 20.0, 30.0]
    for k in (1, 2, 3):
        stopped = _solver(grid, h0)
        _go(stopped, times[k], 10.0, source_rate=src)
        assert hu[k].tobytes() == stopped.download_hu().tobytes()
        assert hv[k].tobytes() == stopped.download_hv().tobytes()
        stopped.close()


def test_dried_inactive_blocks_hold_zero_momentum(grid) -> None:  # noqa: ANN001
    dry_tol = 5e-3
    h0 = sheet(grid)
    on = _solver(grid, h0, dry_tol=dry_tol)
    off = _solver(grid, h0, dry_tol=dry_tol)
    off.active_set = False
    snaps, _, hu, hv = _go(on, 3000.0, 500.0, capture_momentum_snapshots=True)
    snaps_off, _, hu_off, hv_off = _go(off, 3000.0, 500.0, capture_momentum_snapshots=True)
    flowed = np.zeros(on.n_blocks, dtype=bool)
    for h_k, hu_k, hv_k in zip(snaps, hu, hv, strict=True):
        dry = h_k < dry_tol
        assert float(np.abs(hu_k[dry]).max(initial=0.0)) == 0.0
        assert float(np.abs(hv_k[dry]).max(initial=0.0)) == 0.0
        flowed[np.flatnonzero((hu_k != 0.0) | (hv_k != 0.0)) // ACTIVE_BLOCK_CELLS] = True
    inactive = on.download_block_flags()[:, 1] == 0
    assert int(np.count_nonzero(flowed & inactive)) > 0
    for a, b in zip([*snaps, *hu, *hv], [*snaps_off, *hu_off, *hv_off], strict=True):
        np.testing.assert_array_equal(a, b)
    on.close()
    off.close()


def test_a_warm_start_continues_the_run_bit_for_bit(mesh: Path, tmp_path: Path) -> None:
    # Phases of 12 s and 8 s in one run, against 12 s saved to a file and 8 s
    # from it: each phase starts with a prologue and its own output windows,
    # so both take the same steps. Compensation keeps a low word that a warm
    # start does not carry, so both runs are uncompensated here.
    src = [PointSource(2.0, SOURCE_XY, 3.0)]
    full = _workflow(mesh, False, compensated=False)
    whole = full.run([SimulationPhase(12.0, src), SimulationPhase(8.0, src)])
    full.solver.close()
    first = _workflow(mesh, False, compensated=False)
    part = first.run([SimulationPhase(12.0, src)])
    warm = save_warm_start(tmp_path / "warm.npz", part, first.perm, order="original", t_end_s=12.0)
    first.solver.close()
    second = _workflow(mesh, False, compensated=False, initial_state_source=str(warm), initial_state_order="original")
    assert np.array_equal(second.solver.download_h(), part.h_final)
    rest = second.run([SimulationPhase(8.0, src)])
    second.solver.close()
    assert rest.h_final.tobytes() == whole.h_final.tobytes()
    assert rest.hu_final.tobytes() == whole.hu_final.tobytes()
