"""The run loops and their failure paths, against the FP32 oracle (needs an MI355X).

- the phased workflow (sources per phase, output windows, momentum
  snapshots, clamp tally) and the host-driven comparison loops give the same
  bits on the device as the same solver classes driven over the FP32 oracle
  (the loops' context swapped for ``CpuRef``, which takes the same calls);
- the positivity-clamp tally (``INU_FIELD_CLAMP``, reference
  ``resident.py:462-464``) is bitwise the oracle's;
- stale or garbage predictor state in blocks the active set skips is never
  read (reference ``tests/unit/test_gpu_active_set.py:152-167``);
- a non-finite state ends the run early with ``nan_to_num`` snapshots
  (reference ``gpu_resident_batch.py:730-736``), and a CFL bound below one
  tick raises ``RuntimeError`` (``:715-724``).
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.host_driven import SWESolverHipHostDriven
from inundation_amd.hip.solver import SWESolverHipResident, device_mesh_arrays
from inundation_amd.workflow import PointSource, SimulationPhase, SWEWorkflow, WorkflowConfig
from oracle import cpu_ref
from tests import _drive as D
from tests._meshes import MANNING, dam_break, f32_state, valley_grid
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid():  # noqa: ANN201
    return valley_grid(64, 32, 2.0)


@pytest.fixture
def on_oracle(monkeypatch):  # noqa: ANN001, ANN201
    """Within the test, solvers built after ``on_oracle()`` run their loops over the FP32 oracle."""

    def switch() -> None:
        monkeypatch.setattr(SWESolverHipResident, "_open_context", lambda self, device, m: cpu_ref.CpuRef(m))

    return switch


def _same(a, b, what: str) -> None:  # noqa: ANN001
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.dtype == np.float32:
        d = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert d.size == 0, f"{what}: {d.size} words differ, first {d[:5].tolist()}"
    else:
        assert np.array_equal(a, b), what


PHASES = [
    SimulationPhase(7.0, [PointSource(3.0, (20.0, 32.0), 4.0)]),
    SimulationPhase(5.0, [PointSource(8.0, (20.0, 32.0), 4.0), PointSource(1.0, (100.0, 20.0), 0.0)]),
    SimulationPhase(6.0, []),
]


def _workflow_result(grid, compensated: bool):  # noqa: ANN001, ANN202
    wf = SWEWorkflow(WorkflowConfig(mesh_source="unused", manning_n=MANNING, output_interval_s=2.0, dt_max=0.5,
                                    cfl_interval=10, dt_init=0.05, progress=False, capture_momentum_snapshots=True,
                                    track_clamp_mass=True, compensated=compensated, steps_per_graph=16))
    wf.prepare(geometry=grid)
    res = wf.run(PHASES)
    wf.solver.close()
    return res


@pytest.mark.parametrize("compensated", [True, False], ids=["comp", "plain"])
def test_phased_workflow_is_bitwise_the_oracle(grid, on_oracle, compensated: bool) -> None:  # noqa: ANN001
    gpu = _workflow_result(grid, compensated)
    on_oracle()
    ref = _workflow_result(grid, compensated)
    assert gpu.snap_times == ref.snap_times and len(gpu.snapshots) >= 10
    for k, (a, b) in enumerate(zip(gpu.snapshots, ref.snapshots)):
        _same(a, b, f"snapshot {k}")
    for k, (a, b) in enumerate(zip(gpu.hu_snapshots, ref.hu_snapshots)):
        _same(a, b, f"hu snapshot {k}")
    for k, (a, b) in enumerate(zip(gpu.hv_snapshots, ref.hv_snapshots)):
        _same(a, b, f"hv snapshot {k}")
    _same(gpu.h_final, ref.h_final, "h_final")
    _same(gpu.hu_final, ref.hu_final, "hu_final")
    assert gpu.steps_total == ref.steps_total
    assert gpu.volume_final_m3 == ref.volume_final_m3
    assert gpu.clamp_mass_m3 == ref.clamp_mass_m3


def _host_driven(grid, mode: str, source_fn):  # noqa: ANN001, ANN202
    h, hu, hv = f32_state(*dam_break(grid))
    s = SWESolverHipHostDriven(grid, None, h, hu, hv, np.full(grid.N, MANNING, np.float32), dry_tol=1e-3, mode=mode)
    s.deterministic = True  # the oracle sums in the fixed CSR order
    s.compensated = True
    out = s.run(t_end=6.0, output_interval_s=1.5, dt_max=0.25, dt_init=0.02, cfl_interval=7, progress=False,
                source_fn=source_fn, capture_momentum_snapshots=mode == "baseline",
                output_every_steps=None)
    steps = s.steps_total
    s.close()
    return out, steps


def _rain(grid):  # noqa: ANN001, ANN202
    def fn(t: float) -> np.ndarray:
        return (grid.area * (1e-4 * (1.0 + np.sin(t)))).astype(np.float32)

    return fn


@pytest.mark.parametrize("mode,sources", [("baseline", False), ("baseline", True), ("fixed_dt_batch_barrier", False)])
def test_host_driven_loops_are_bitwise_the_oracle(grid, on_oracle, mode: str, sources: bool) -> None:  # noqa: ANN001
    fn = _rain(grid) if sources else None
    gpu, n_gpu = _host_driven(grid, mode, fn)
    on_oracle()
    ref, n_ref = _host_driven(grid, mode, fn)
    assert n_gpu == n_ref and gpu[1] == ref[1]
    for part, (xs, ys) in enumerate(zip(gpu, ref)):
        if part == 1:
            continue
        for k, (a, b) in enumerate(zip(xs, ys)):
            _same(a, b, f"{mode}: output {part}, snapshot {k}")


@pytest.fixture(scope="module")
def ctx():  # noqa: ANN201
    c = gpu_context()
    yield c
    c.close()


def _mesh(ctx, grid):  # noqa: ANN001, ANN202
    m = device_mesh_arrays(grid)
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    return m


def _rough_film(grid):  # noqa: ANN001, ANN202
    """A film up to 1 cm deep with random momentum: at CFL 2 some cells overshoot below zero and clamp."""
    rng = np.random.default_rng(3)
    h = rng.uniform(0.0, 0.01, grid.N).astype(np.float32)
    return h, (h * rng.normal(0.0, 2.0, grid.N)).astype(np.float32), (h * rng.normal(0.0, 2.0, grid.N)).astype(np.float32)


@pytest.mark.parametrize("case,cfl", [("rough_film", 2.0), ("dam_break", 0.8)])
def test_clamp_tally_is_bitwise_the_oracle(ctx, grid, case: str, cfl: float) -> None:  # noqa: ANN001
    m = _mesh(ctx, grid)
    ref = cpu_ref.CpuRef(m)
    h, hu, hv = _rough_film(grid) if case == "rough_film" else f32_state(*dam_break(grid))
    flags = D.FLAGS_DEFAULT | L.FLAG_TRACK_CLAMP
    for be in (ctx, ref):
        D.setup(be, grid.N, h, hu, hv, MANNING, dry_tol=1e-3, cfl=cfl, flags=flags)
        be.upload(L.FIELD_CLAMP, np.zeros(grid.N, np.float32))
        D.start(be, 30.0, 0.5, 0.01)
    ctx.steps(60, use_graph=True)
    ref.steps(60)
    got, want = ctx.download(L.FIELD_CLAMP, grid.N), ref.download(L.FIELD_CLAMP, grid.N)
    if case == "rough_film":
        assert np.count_nonzero(want) > 0, "the case must clamp somewhere"
    _same(got, want, "clamp tally")
    D.assert_same(D.snapshot(ctx, grid.N, -(-grid.N // 256)), D.snapshot(ref, grid.N, -(-grid.N // 256)), "state")


def test_garbage_predictor_state_in_skipped_blocks_is_ignored() -> None:
    grid = valley_grid(256, 64, 2.0)  # 64 blocks: the dam's two-ring leaves most of them skipped
    m = device_mesh_arrays(grid)
    h, hu, hv = f32_state(*dam_break(grid))
    garbage = np.full(grid.N, 0.5, dtype=np.float32)
    nb = -(-grid.N // 256)
    ref = cpu_ref.CpuRef(m)
    on, off = gpu_context(), gpu_context()
    try:
        for be, flags in ((on, D.FLAGS_DEFAULT), (off, L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED),
                          (ref, D.FLAGS_DEFAULT)):
            if be is not ref:
                be.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"],
                            m["cidx"])
            D.setup(be, grid.N, h, hu, hv, MANNING, dry_tol=1e-3, flags=flags)
            if be is not ref:
                for f in (L.FIELD_H1, L.FIELD_HU1, L.FIELD_HV1, L.FIELD_H1_LO):
                    be.upload(f, garbage)
            D.start(be, 30.0, 0.5, 0.01)
        assert on.download(L.FIELD_BLOCK_ACTIVE, nb, np.uint32).sum() < nb, "some block must be skipped"
        for be in (on, off):
            be.steps(40, use_graph=True)
        ref.steps(40)
        D.assert_same(D.snapshot(on, grid.N, nb), D.snapshot(ref, grid.N, nb), "active set on vs oracle")
        a, b = D.snapshot(on, grid.N, nb), D.snapshot(off, grid.N, nb)
        for k in ("h", "hu", "hv", "h_lo", "ctrl", "dt"):
            _same(a[k], b[k], f"on vs off: {k}")
    finally:
        on.close()
        off.close()


def test_a_non_finite_state_ends_the_run_with_nan_to_num_snapshots(grid) -> None:  # noqa: ANN001
    h, hu, hv = f32_state(*dam_break(grid))
    h = h.copy()
    h[int(np.flatnonzero(h > 0)[0])] = np.nan
    s = SWESolverHipResident(grid, None, h, hu, hv, np.full(grid.N, MANNING, np.float32), dry_tol=1e-3)
    try:
        snaps, times = s.run(t_end=20.0, output_interval_s=5.0, dt_max=0.5, dt_init=0.01, cfl_interval=1,
                             progress=False)
    finally:
        s.close()
    assert times == [0.0, 5.0], times
    assert np.isnan(snaps[0]).sum() == 1  # the initial state as given
    assert np.isfinite(snaps[-1]).all()


def test_a_cfl_bound_below_one_tick_raises(grid) -> None:  # noqa: ANN001
    h, hu, hv = f32_state(*dam_break(grid))
    s = SWESolverHipResident(grid, None, h, hu, hv, np.full(grid.N, MANNING, np.float32), dry_tol=1e-3, cfl=1e-12)
    try:
        with pytest.raises(RuntimeError, match="no time progress"):
            s.run(t_end=5.0, output_interval_s=1.0, dt_max=0.5, dt_init=0.01, cfl_interval=1, progress=False)
    finally:
        s.close()
