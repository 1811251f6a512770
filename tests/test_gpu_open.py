"""Open boundaries on the HIP backend against the FP32 oracle (needs an MI355X).

- the ghost-state and boundary-flux probe bit for bit against the oracle's
  (the oracle's formulas are pinned to closed forms in test_oracle_open.py);
- channel runs bit for bit against the oracle: state, low words, block flags,
  clock, dt words and the stretch tally, for inflow, stage, normal depth, free
  outflow, the wall kind and a film-thin inflow, with both arithmetic orders,
  with the active set on and off, and across a change of conditions between
  two runs (reference ``tests/unit/test_gpu_open_boundaries.py``);
- the workflow: volumes line up with the snapshots, the balance closes, and
  closing a stretch releases its blocks.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays, mark_open_edges
from inundation_amd.mesh.boundaries import select_boundary_edges
from oracle import cpu_ref
from tests import _drive as D
from tests._meshes import channel_end, channel_grid
from tests.conftest import gpu_context
from tests.test_oracle_open import (
    DRY,
    FREE,
    G,
    INFLOW,
    MANNING,
    NORMAL,
    SLOPE,
    STAGE,
    WALL,
    _conditions,
    _run,
)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():  # noqa: ANN201
    c = gpu_context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def channel():  # noqa: ANN201
    g = channel_grid()
    ends = [select_boundary_edges(g, *channel_end(x)) for x in (0.0, 3000.0)]
    m = device_mesh_arrays(g)
    ptr = mark_open_edges(g, m, ends)
    return g, ends, m, ptr


def test_open_probe_is_bitwise_the_oracle(ctx) -> None:  # noqa: ANN001
    rng = np.random.default_rng(11)
    n = 1 << 16
    h = np.concatenate([
        np.zeros(n // 8),
        rng.uniform(0.0, 2e-4, n // 8),
        rng.uniform(0.0, 0.02, n // 4),
        10.0 ** rng.uniform(-3, 1.3, n // 2),
    ])
    ang = rng.uniform(0, 2 * np.pi, n)
    cases = np.column_stack([
        h,
        h * rng.normal(0, 2.0, n),
        h * rng.normal(0, 2.0, n),
        np.cos(ang),
        np.sin(ang),
        rng.normal(0, 1.0, n),
    ]).astype(np.float32)
    kinds = rng.integers(0, 5, n).astype(np.uint32)
    scale = np.array([1.0, 1.0, 2.0, 5.0, 20.0])[kinds]
    values = (scale * rng.uniform(0.0, 1.0, n)).astype(np.float32)
    values[rng.uniform(size=n) < 0.05] = 0.0
    got = ctx.probe_open(cases, kinds, values, G, DRY)
    want = cpu_ref.probe_open(cases, kinds, values, G, DRY)
    assert np.isfinite(want).all()
    bad = np.flatnonzero((got.view(np.uint32) != want.view(np.uint32)).any(axis=1))
    assert bad.size == 0, f"{bad.size} cases differ; first {bad[:4].tolist()}: {got[bad[:2]]} vs {want[bad[:2]]}"


def _volumes(be, n: int) -> np.ndarray:  # noqa: ANN001
    return be.open_volumes(n) if isinstance(be, L.Context) else be.open_volumes()


def _snap(be, g, nb: int) -> dict[str, np.ndarray]:  # noqa: ANN001
    out = D.snapshot(be, g.N, nb)
    out["forced"] = be.download(L.FIELD_BLOCK_FORCED, nb, np.uint32)
    out["volumes"] = _volumes(be, 2)
    return out


CASES = {
    # name: initial depth, initial x-velocity, [(up, down, t0, t1), ...]
    "inflow_into_dry": (0.0, 0.0, [((INFLOW, 50.0), (NORMAL, SLOPE), 0.0, 60.0)]),
    "stage_and_free": (0.5, -0.5, [((STAGE, 10.4), (FREE, 0.0), 0.0, 30.0)]),
    "wall_kind_and_drawing_outlet": (0.5, -2.0, [((WALL, 0.0), (NORMAL, 1e-5), 0.0, 5.0)]),
    "film_inflow": (0.0, 0.0, [((INFLOW, 1e-3), (WALL, 0.0), 0.0, 30.0)]),
    "two_phases": (
        0.0,
        0.0,
        [((INFLOW, 50.0), (WALL, 0.0), 0.0, 12.0), ((WALL, 0.0), (FREE, 0.0), 12.0, 20.0)],
    ),
}
ORDERS = {
    "fast": D.FLAGS_DEFAULT,
    "exact": D.FLAGS_DEFAULT | 16,
    "all_blocks": D.FLAGS_DEFAULT & ~L.FLAG_ACTIVE_SET,
    "uncompensated": D.FLAGS_DEFAULT & ~L.FLAG_COMPENSATED,
}


@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("case", list(CASES))
def test_channel_runs_are_bitwise_the_oracle(ctx, channel, case: str, order: str) -> None:  # noqa: ANN001
    g, ends, m, ptr = channel
    depth, u0, phases = CASES[case]
    flags = ORDERS[order]
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    ref = cpu_ref.CpuRef(m)
    h0 = np.full(g.N, depth, dtype=np.float32)
    hu0 = (u0 * h0).astype(np.float32)
    z = np.zeros(g.N, dtype=np.float32)
    for be in (ctx, ref):
        D.setup(be, g.N, h0, hu0, z, MANNING, g=G, dry_tol=DRY, flags=flags)
        be.set_open_stretches(ptr)
    for k, (up, down, t0, t1) in enumerate(phases):
        kinds, values = _conditions(g, ends, up, down)
        steps = []
        for be in (ctx, ref):
            be.set_open_conditions(kinds, values)
            steps.append(_run(be, t0, t1))
        assert steps[0] == steps[1] and steps[0] > 0
        D.assert_same(_snap(ctx, g, ref.NB), _snap(ref, g, ref.NB), f"{case}/{order} phase {k}")
    vol = _volumes(ctx, 2)
    if case == "inflow_into_dry":
        assert vol[0, 1] == pytest.approx(50.0 * 60.0, rel=1e-6)
    if case == "wall_kind_and_drawing_outlet":
        assert not vol.any()


def test_stretches_split_the_tally(ctx, channel) -> None:  # noqa: ANN001
    g, ends, m, _ = channel
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    ref = cpu_ref.CpuRef(m)
    four = np.array([0, 10, 20, 30, 40], dtype=np.uint32)
    h0 = np.full(g.N, 0.5, dtype=np.float32)
    z = np.zeros(g.N, dtype=np.float32)
    kinds, values = _conditions(g, ends, (INFLOW, 30.0), (NORMAL, SLOPE))
    for be in (ctx, ref):
        D.setup(be, g.N, h0, z, z, MANNING, g=G, dry_tol=DRY)
        be.set_open_stretches(four)
        be.set_open_conditions(kinds, values)
        _run(be, 0.0, 20.0)
    got, want = ctx.open_volumes(4), ref.open_volumes()
    assert np.array_equal(got, want)
    assert got[:2, 1].sum() == pytest.approx(30.0 * 20.0, rel=1e-6)
    assert (got[2:, 0] > 0).all()
    with pytest.raises(ValueError):
        ctx.open_volumes(2)


# ---- the workflow -----------------------------------------------------------


@pytest.fixture(scope="module")
def channel_file(tmp_path_factory):  # noqa: ANN001, ANN201
    """The channel as a mesh file, as the reference's tests write it."""
    from inundation_amd.mesh.loader import write_mesh_parquet
    from inundation_amd.mesh.synth import quad_grid

    verts, faces = quad_grid(300, 20, 10.0)
    cx = (np.arange(300 * 20) % 300 + 0.5) * 10.0
    path = tmp_path_factory.mktemp("open") / "channel.parquet"
    write_mesh_parquet(path, verts, faces, (10.0 - SLOPE * cx).astype(np.float32))
    return str(path)


def _workflow(path: str, initial_depth: float | None = None, **cfg):  # noqa: ANN001, ANN202
    from inundation_amd.workflow import OpenBoundary, SWEWorkflow, WorkflowConfig

    state = None
    if initial_depth is not None:
        h = np.full(300 * 20, initial_depth, dtype=np.float32)
        state = {"h": h, "hu": np.zeros_like(h), "hv": np.zeros_like(h)}
    wf = SWEWorkflow(
        WorkflowConfig(
            mesh_source=path,
            manning_n=MANNING,
            output_interval_s=5.0,
            dt_max=1.0,
            cfl_interval=1,
            progress=False,
            track_clamp_mass=True,
            dry_tol=DRY,
            initial_state_source=state,
            initial_state_order="original",
            open_boundaries=[
                OpenBoundary("up", *channel_end(0.0)),
                OpenBoundary("down", *channel_end(3000.0)),
            ],
            **cfg,
        )
    )
    wf.prepare()
    return wf


def test_workflow_volumes_line_up_with_the_snapshots(channel_file) -> None:  # noqa: ANN001
    from inundation_amd.workflow import Inflow, SimulationPhase, Wall

    wf = _workflow(channel_file)
    res = wf.run([
        SimulationPhase(12.0, [], {"up": Inflow(50.0, SLOPE), "down": Wall()}),
        SimulationPhase(8.0, [], {"up": Inflow(20.0, SLOPE), "down": Wall()}),
    ])
    v = res.boundary_volumes_m3["up"]
    assert v.shape == (len(res.snap_times), 2)
    assert (np.diff(v[:, 1]) >= 0.0).all()
    t = np.asarray(res.snap_times)
    for k in np.flatnonzero(t == 12.0):
        assert v[k, 1] == pytest.approx(50.0 * 12.0, rel=1e-6)
    assert v[-1, 1] == pytest.approx(50.0 * 12.0 + 20.0 * 8.0, rel=1e-6)
    assert res.volume_final_m3 == pytest.approx(v[-1, 1], rel=1e-6)
    wf.solver.close()


def test_workflow_balance_closes_while_water_enters_and_leaves(channel_file) -> None:  # noqa: ANN001
    from inundation_amd.workflow import Inflow, NormalDepth, SimulationPhase

    wf = _workflow(channel_file, initial_depth=0.5)
    start = 0.5 * float(wf.geom.area.astype(np.float64).sum())
    res = wf.run([SimulationPhase(120.0, [], {"up": Inflow(50.0, SLOPE), "down": NormalDepth(SLOPE)})])
    went_in = res.boundary_volumes_m3["up"][-1, 1]
    went_out = res.boundary_volumes_m3["down"][-1, 0]
    assert went_out > 0.0
    assert abs(start + went_in - went_out - res.volume_final_m3) / start < 1e-7
    wf.solver.close()


def test_closing_a_stretch_releases_its_blocks(channel_file) -> None:  # noqa: ANN001
    from inundation_amd.workflow import Inflow, SimulationPhase, Wall

    wf = _workflow(channel_file)
    wf.run([SimulationPhase(5.0, [], {"up": Inflow(50.0, SLOPE), "down": Wall()})])
    assert wf.solver.download_block_flags()[:, 2].sum() > 0
    wf.run([SimulationPhase(5.0, [], {"up": Wall(), "down": Wall()})])
    assert wf.solver.download_block_flags()[:, 2].sum() == 0
    wf.solver.close()


def test_repeat_runs_are_bit_identical(channel_file) -> None:  # noqa: ANN001
    from inundation_amd.workflow import Inflow, SimulationPhase, Stage

    n_cells = 300 * 20
    phases = [SimulationPhase(30.0, [], {"up": Inflow(50.0, SLOPE), "down": Stage(7.3)})]
    bits = []
    for _ in range(2):
        wf = _workflow(channel_file)
        r = wf.run(phases)
        bits.append(r.h_final.tobytes() + b"".join(v.tobytes() for v in r.boundary_volumes_m3.values()))
        wf.solver.close()
    assert bits[0] == bits[1]
    assert math.isfinite(float(np.frombuffer(bits[0][: 4 * n_cells], np.float32).sum()))
