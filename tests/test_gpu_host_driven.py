"""Host-driven comparison backends on the HIP kernels (needs an MI355X).

``baseline`` and ``fixed_dt_batch_barrier`` keep the reference's host loops
(CFL dt estimated every ``cfl_interval`` steps, 0.8 of it kept fixed in
between; ``vulkan/solver.py:499-760``, ``fixed_dt_batch_barrier.py:68-270``).

- Against the resident solver at the same CFL number they agree to the
  scheme's truncation error: momentum within 1 % relative L2 at each output
  time (the reference's own A/B bound, ``test_gpu_momentum_snapshots.py:287-330``);
- the fixed-dt batches land on the output times, or snapshot every
  ``output_every_steps`` steps; the baseline snapshots on the first step at
  or past each output time;
- sources arrive whole, also a time-varying ``source_fn``;
- ``fixed_dt_batch_barrier`` refuses momentum snapshots and open
  boundaries, as the reference does; the workflow runs ``baseline`` with
  open boundaries and its tally closes the balance.
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip.host_driven import SWESolverHipHostDriven
from inundation_amd.hip.solver import SWESolverHipResident
from inundation_amd.tuning import CFL_HOST_DRIVEN
from tests._meshes import MANNING, valley_grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid():  # noqa: ANN201
    return valley_grid(256, 128, 2.0)


def _source(grid, q: float = 1.0) -> np.ndarray:  # noqa: ANN001
    src = np.zeros(grid.N, dtype=np.float32)
    d = (grid.centroid[:, 0] - 256 * 2.0 * 0.2) ** 2 + (grid.centroid[:, 1] - 128 * 2.0 / 2) ** 2
    src[int(np.argmin(d))] = q
    return src


def _make(cls, grid, **kw):  # noqa: ANN001, ANN003, ANN202
    z = np.zeros(grid.N, dtype=np.float32)
    return cls(grid, None, z, z, z, np.full(grid.N, MANNING, np.float32), **kw)


def _go(s, t_end: float, interval: float, **kw):  # noqa: ANN001, ANN003, ANN202
    kw.setdefault("cfl_interval", 10)
    return s.run(t_end=t_end, output_interval_s=interval, dt_max=1.0, dt_init=0.01, progress=False, **kw)


@pytest.mark.parametrize("mode", ["baseline", "fixed_dt_batch_barrier"])
def test_host_driven_momentum_matches_the_resident_solver(grid, mode: str) -> None:  # noqa: ANN001
    src = _source(grid)
    res = _make(SWESolverHipResident, grid, cfl=CFL_HOST_DRIVEN)
    _, times, hu, hv = _go(res, 60.0, 20.0, source_rate=src, capture_momentum_snapshots=True)
    res.close()
    assert times == [0.0, 20.0, 40.0, 60.0]
    host = _make(SWESolverHipHostDriven, grid, mode=mode)
    for k in (1, 2, 3):
        _go(host, times[k], 20.0, source_rate=src, cfl_interval=1, resume=k > 1, t_start=times[k - 1])
        q_ref = np.concatenate([hu[k], hv[k]]).astype(np.float64)
        q = np.concatenate([host.download_hu(), host.download_hv()]).astype(np.float64)
        assert np.linalg.norm(q_ref) > 0.0
        assert np.linalg.norm(q - q_ref) / np.linalg.norm(q_ref) < 0.01
    host.close()


def test_fixed_dt_batches_land_on_output_times_and_sources_arrive(grid) -> None:  # noqa: ANN001
    host = _make(SWESolverHipHostDriven, grid, mode="fixed_dt_batch_barrier")
    snaps, times = _go(host, 30.0, 7.0, source_rate=_source(grid, 2.0), cfl_interval=25)
    assert times == [0.0, 7.0, 14.0, 21.0, 28.0, 30.0]
    area = grid.area.astype(np.float64)
    vol = float((snaps[-1].astype(np.float64) * area).sum())
    assert vol == pytest.approx(60.0, rel=1e-5)
    with pytest.raises(ValueError, match="momentum"):
        _go(host, 31.0, 1.0, capture_momentum_snapshots=True)
    host.close()


def test_fixed_dt_step_cadence(grid) -> None:  # noqa: ANN001
    host = _make(SWESolverHipHostDriven, grid, mode="fixed_dt_batch_barrier")
    _, times = _go(host, 20.0, 5.0, source_rate=_source(grid), cfl_interval=7, output_every_steps=12)
    assert len(times) >= 3 and times[0] == 0.0 and times[-1] == 20.0
    assert host.steps_total >= 12 * (len(times) - 2)
    host.close()


def test_baseline_snapshots_at_or_past_output_times_with_a_time_varying_source(grid) -> None:  # noqa: ANN001
    host = _make(SWESolverHipHostDriven, grid, mode="baseline")
    cell = int(np.flatnonzero(_source(grid))[0])

    def rate(t: float) -> np.ndarray:
        r = np.zeros(grid.N, dtype=np.float32)
        r[cell] = 1.0 + 0.1 * t
        return r

    snaps, times, hu, hv = _go(host, 10.0, 3.0, source_fn=rate, capture_momentum_snapshots=True, cfl_interval=5)
    assert len(snaps) == len(times) == len(hu) == len(hv)
    for k, t_out in enumerate((3.0, 6.0, 9.0), start=1):
        assert t_out <= times[k] < t_out + 1.0
    assert times[-1] == 10.0
    vol = float((snaps[-1].astype(np.float64) * grid.area.astype(np.float64)).sum())
    # The rate is sampled at each step's start, dt <= 1 s: the integral
    # 10 + 0.05 t^2 = 15 m^3 less at most 0.05 * 10 s * 1 s.
    assert 15.0 - 0.5 - 1e-3 <= vol <= 15.0 + 1e-3
    host.close()


def test_workflow_runs_the_baseline_with_open_boundaries(tmp_path) -> None:  # noqa: ANN001
    from inundation_amd.mesh.loader import write_mesh_parquet
    from inundation_amd.mesh.synth import quad_grid
    from inundation_amd.workflow import (
        Inflow,
        NormalDepth,
        OpenBoundary,
        SimulationPhase,
        SWEWorkflow,
        WorkflowConfig,
    )

    verts, faces = quad_grid(120, 10, 10.0)
    cx = (np.arange(1200) % 120 + 0.5) * 10.0
    path = tmp_path / "channel.parquet"
    write_mesh_parquet(path, verts, faces, (10.0 - 1e-3 * cx).astype(np.float32))
    h = np.full(1200, 0.5, dtype=np.float32)
    wf = SWEWorkflow(
        WorkflowConfig(
            mesh_source=str(path),
            manning_n=MANNING,
            output_interval_s=10.0,
            dt_max=1.0,
            cfl_interval=5,
            progress=False,
            dry_tol=1e-4,
            solver_impl="baseline",
            initial_state_source={"h": h, "hu": np.zeros_like(h), "hv": np.zeros_like(h)},
            initial_state_order="original",
            open_boundaries=[
                OpenBoundary("up", [(0.0, 0.0), (0.0, 100.0)], 2.5),
                OpenBoundary("down", [(1200.0, 0.0), (1200.0, 100.0)], 2.5),
            ],
        )
    )
    wf.prepare()
    assert isinstance(wf.solver, SWESolverHipHostDriven)
    res = wf.run([SimulationPhase(60.0, [], {"up": Inflow(10.0, 1e-3), "down": NormalDepth(1e-3)})])
    v = res.boundary_volumes_m3
    assert v["up"].shape == (len(res.snap_times), 2)
    start = 0.5 * 1200 * 100.0
    went_in, went_out = v["up"][-1, 1], v["down"][-1, 0]
    assert went_in == pytest.approx(600.0, rel=1e-5)
    assert abs(start + went_in - went_out - res.volume_final_m3) / start < 1e-6
    wf.solver.close()
