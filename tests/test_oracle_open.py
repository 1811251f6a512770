"""Open boundaries in the FP32 oracle: the ghost formulas against closed forms, then channel runs.

The ghost and boundary-flux cases are the reference's own known answers
(``tests/unit/test_gpu_open_ghost.py:137-205``: free outflow, normal depth,
stage, dam break over a dry cell, the wall kind, the inflow root found by
bisection, a rotated normal). The channel runs check what the formulas alone
cannot: a stage at the still-water level keeps a sloping lake bitwise still,
the stretch tally closes the volume balance, an inflow's volume arrives whole
even as a film, outflow conditions draw nothing in, an outlet that would draw
water in is a wall, and the active set changes nothing
(``tests/unit/test_gpu_open_boundaries.py:108-270``).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays, mark_open_edges
from inundation_amd.mesh.boundaries import conveyance_split, select_boundary_edges
from inundation_amd.tuning import dt_to_ticks
from oracle import cpu_ref
from tests._drive import FLAGS_DEFAULT
from tests._meshes import channel_end, channel_grid

G = 9.81
DRY = 1e-4
SLOPE = 1e-3
MANNING = 0.03
WALL, FREE, NORMAL, STAGE, INFLOW = L.OPEN_WALL, L.OPEN_FREE, L.OPEN_NORMAL, L.OPEN_STAGE, L.OPEN_INFLOW


def _ghost(h, hu, hv, kind, value=0.0, normal=(1.0, 0.0), fh=0.0):  # noqa: ANN001, ANN202
    cases = np.array([[h, hu, hv, normal[0], normal[1], fh]], dtype=np.float32)
    return tuple(float(x) for x in cpu_ref.probe_open(cases, [kind], [value], G, DRY)[0])


def _root(q: float, riemann: float) -> float:
    """hG solving -q/hG + 2 sqrt(g hG) = riemann, by bisection on sqrt(hG)."""
    rg = math.sqrt(G)
    lo, hi = 1e-9, 100.0
    for _ in range(200):
        s = 0.5 * (lo + hi)
        if 2.0 * rg * s - q / (s * s) < riemann:
            lo = s
        else:
            hi = s
    return hi * hi


# ---- ghost states and boundary mass flux ----------------------------------


def test_free_outflow_copies_the_interior() -> None:
    assert _ghost(1.0, 0.5, 0.2, FREE, fh=0.5) == pytest.approx((1.0, 0.5, 0.2, 0.5), rel=1e-6)


def test_free_outflow_mirrors_flow_running_inward() -> None:
    assert _ghost(1.0, -0.5, 0.2, FREE, fh=-0.5) == pytest.approx((1.0, 0.5, 0.2, 0.0), rel=1e-6)


@pytest.mark.parametrize("kind", [FREE, NORMAL])
def test_outflow_conditions_never_take_a_negative_flux(kind: int) -> None:
    assert _ghost(1.0, 0.0, 0.0, kind, 1.0, fh=-0.25)[3] == 0.0


def test_normal_depth_uses_the_manning_velocity() -> None:
    value = math.sqrt(1e-3) / 0.03
    h = 0.8
    un = value * h ** (2.0 / 3.0)
    assert _ghost(h, 0.1, 0.05, NORMAL, value) == pytest.approx((h, h * un, 0.05, 0.0), rel=1e-5)


def test_a_stage_keeps_the_outgoing_characteristic() -> None:
    h, u, depth = 1.0, 0.5, 2.0
    un = u + 2.0 * (math.sqrt(G * h) - math.sqrt(G * depth))
    assert _ghost(h, h * u, 0.0, STAGE, depth, fh=-0.7) == pytest.approx((depth, depth * un, 0.0, -0.7), rel=1e-5)


def test_a_stage_cannot_reach_a_supercritical_outflow() -> None:
    h, u = 0.1, 3.0
    assert _ghost(h, h * u, 0.0, STAGE, 2.0)[:3] == pytest.approx((h, h * u, 0.0), rel=1e-5)


def test_a_stage_over_a_dry_cell_enters_as_a_dam_break() -> None:
    depth = 0.5
    un = -2.0 * math.sqrt(G * depth)
    assert _ghost(0.0, 0.0, 0.0, STAGE, depth)[:2] == pytest.approx((depth, depth * un), rel=1e-5)


def test_a_wall_kind_mirrors_and_passes_nothing() -> None:
    assert _ghost(1.0, 0.5, 0.2, WALL, fh=0.3) == pytest.approx((1.0, -0.5, 0.2, 0.0), rel=1e-6)


@pytest.mark.parametrize(
    ("h", "u", "q"), [(0.0, 0.0, 0.25), (0.05, 0.0, 2.0), (1.0, -0.2, 0.25), (3.0, 0.5, 5.0), (0.2, 0.0, 17.0)]
)
def test_inflow_solves_the_outgoing_characteristic(h: float, u: float, q: float) -> None:
    depth = _root(q, u + 2.0 * math.sqrt(G * h))
    assert _ghost(h, h * u, 0.7 * h, INFLOW, q, fh=9.0) == pytest.approx((depth, -q, 0.0, -q), rel=1e-4, abs=1e-6)


def test_ghosts_follow_the_edge_normal() -> None:
    nx, ny, h, q = 0.6, 0.8, 1.0, 0.5
    depth = _root(q, 2.0 * math.sqrt(G * h))
    assert _ghost(h, 0.0, 0.0, INFLOW, q, normal=(nx, ny))[:3] == pytest.approx((depth, -q * nx, -q * ny), rel=1e-4)


def test_an_inflow_of_nothing_is_a_wall_mirror() -> None:
    assert _ghost(1.0, 0.5, 0.2, INFLOW, 0.0)[:3] == pytest.approx((1.0, -0.5, 0.2), rel=1e-6)


# ---- stretch selection and conveyance split --------------------------------


def test_conveyance_split_carries_the_discharge_along_the_low_part() -> None:
    lengths = np.full(6, 10.0, dtype=np.float32)
    beds = np.array([3.0, 2.0, 1.0, 1.0, 2.0, 3.0], dtype=np.float32)
    n = np.full(6, 0.03, dtype=np.float32)
    q = conveyance_split(40.0, lengths, beds, n, 1e-3)
    assert q.dtype == np.float32
    assert float((q.astype(np.float64) * lengths).sum()) == pytest.approx(40.0, rel=1e-7)
    # The level stays below the top edges, and the section is symmetric.
    assert q[0] == q[5] == 0.0
    assert q[2] == pytest.approx(q[3], rel=1e-6) and q[1] == pytest.approx(q[4], rel=1e-6)
    # One level L for all edges: q len = len sqrt(S) / n (L - bed)^(5/3).
    k = math.sqrt(1e-3) / 0.03
    level = 1.0 + (float(q[2]) / k) ** 0.6
    assert float(q[1]) == pytest.approx(k * (level - 2.0) ** (5.0 / 3.0), rel=1e-5)
    # A flat section splits evenly.
    flat = conveyance_split(40.0, lengths[:2], np.zeros(2, np.float32), n[:2], 1e-3)
    assert flat == pytest.approx([2.0, 2.0], rel=1e-7)
    assert not conveyance_split(0.0, lengths, beds, n, 1e-3).any()


def test_a_line_across_the_channel_picks_its_end_edges() -> None:
    g = channel_grid()
    for x, nx in ((0.0, -1.0), (3000.0, 1.0)):
        line, tol = channel_end(x)
        e = select_boundary_edges(g, line, tol)
        assert e.size == 20
        assert (g.edge_cellR[e] == -1).all() and (g.edge_nx[e] == nx).all()


# ---- channel runs ----------------------------------------------------------


@pytest.fixture(scope="module")
def channel():  # noqa: ANN201
    g = channel_grid()
    ends = [select_boundary_edges(g, *channel_end(x)) for x in (0.0, 3000.0)]
    return g, ends


def _oracle(g, ends, flags: int = FLAGS_DEFAULT):  # noqa: ANN001, ANN202
    m = device_mesh_arrays(g)
    ptr = mark_open_edges(g, m, ends)
    ref = cpu_ref.CpuRef(m)
    ref.set_params(G, DRY, 0.8, flags)
    ref.set_manning(np.full(g.N, MANNING, dtype=np.float32))
    ref.set_sources(np.empty(0, np.uint32), np.empty(0, np.float32))
    ref.set_open_stretches(ptr)
    return ref


def _conditions(g, ends, up, down):  # noqa: ANN001, ANN202
    """(kind, value) per end, as values per edge: value is a scalar or 'stage level' / 'inflow Q'."""
    kinds, values = [], []
    for edges, (kind, value) in zip(ends, (up, down), strict=True):
        cells = g.edge_cellL[edges]
        if kind == STAGE:
            v = np.maximum(value - g.zb[cells].astype(np.float64), 0.0).astype(np.float32)
        elif kind == INFLOW:
            v = conveyance_split(value, g.edge_len[edges], g.zb[cells], np.full(edges.size, MANNING, np.float32), SLOPE)
        elif kind == NORMAL:
            v = np.full(edges.size, math.sqrt(value) / MANNING, dtype=np.float32)
        else:
            v = np.zeros(edges.size, dtype=np.float32)
        kinds.append(np.full(edges.size, kind, dtype=np.uint32))
        values.append(v)
    return np.concatenate(kinds), np.concatenate(values)


def _run(be, t0: float, t1: float, dt_max: float = 1.0, dt_init: float = 1e-2) -> int:  # noqa: ANN001
    """One run window [t0, t1] as the solver drives it: control block, prologue, step blocks."""
    c = L.InuCtrl()
    c.now_ticks = round(t0 * 2**20)
    c.stop_ticks = round(t1 * 2**20)
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(dt_max)
    c.dry_dt_ticks = dt_to_ticks(min(dt_init, dt_max))
    c.cfl_safety = 1.0
    be.ctrl_write(c)
    be.prologue()
    c = be.ctrl_read()
    while int(c.now_ticks) < int(c.stop_ticks):
        before = int(c.now_ticks)
        be.steps(32)
        c = be.ctrl_read()
        assert int(c.now_ticks) > before
    return int(c.steps)


def _volume(be, g) -> float:  # noqa: ANN001
    h = be.download(L.FIELD_H, g.N).astype(np.float64) + be.download(L.FIELD_H_LO, g.N).astype(np.float64)
    return float((h * g.area.astype(np.float64)).sum())


def test_a_stage_at_the_lake_level_keeps_the_lake_at_rest(channel) -> None:  # noqa: ANN001
    # The float32 depths put the level within an ulp of 10.4 m, so the lake is
    # at rest to rounding, as it is between walls.
    g, ends = channel
    ref = _oracle(g, ends)
    h0 = np.maximum(10.4 - g.zb.astype(np.float64), 0.0).astype(np.float32)
    z = np.zeros(g.N, dtype=np.float32)
    ref.set_state(h0, z, z)
    ref.set_open_conditions(*_conditions(g, ends, (STAGE, 10.4), (STAGE, 10.4)))
    _run(ref, 0.0, 300.0)
    h = ref.download(L.FIELD_H, g.N)
    speed = np.hypot(ref.download(L.FIELD_HU, g.N), ref.download(L.FIELD_HV, g.N)) / h
    assert float(speed.max()) < 1e-6
    assert float(np.abs(h - h0).max()) < 1e-6
    assert float(np.abs(ref.open_volumes()).max()) < 1e-3


def test_the_tally_closes_the_balance_while_water_enters_and_leaves(channel) -> None:  # noqa: ANN001
    g, ends = channel
    ref = _oracle(g, ends)
    h0 = np.full(g.N, 0.5, dtype=np.float32)
    z = np.zeros(g.N, dtype=np.float32)
    ref.set_state(h0, z, z)
    start = _volume(ref, g)
    ref.set_open_conditions(*_conditions(g, ends, (INFLOW, 50.0), (NORMAL, SLOPE)))
    _run(ref, 0.0, 120.0)
    v = ref.open_volumes()
    went_in, went_out = v[0, 1], v[1, 0]
    assert went_out > 0.0 and v[0, 0] == 0.0 and v[1, 1] == 0.0
    assert went_in == pytest.approx(50.0 * 120.0, rel=1e-6)
    assert abs(start + went_in - went_out - _volume(ref, g)) / start < 1e-7


def test_a_trickle_too_thin_to_wet_a_cell_still_arrives(channel) -> None:  # noqa: ANN001
    g, ends = channel
    ref = _oracle(g, ends)
    z = np.zeros(g.N, dtype=np.float32)
    ref.set_state(z, z, z)
    ref.set_open_conditions(*_conditions(g, ends, (INFLOW, 1e-3), (WALL, 0.0)))
    _run(ref, 0.0, 60.0)
    assert float(ref.download(L.FIELD_H, g.N).max()) < DRY
    assert ref.open_volumes()[0, 1] == pytest.approx(1e-3 * 60.0, rel=1e-6)
    assert _volume(ref, g) == pytest.approx(1e-3 * 60.0, rel=1e-6)


@pytest.mark.parametrize("down", [(FREE, 0.0), (NORMAL, SLOPE)], ids=["free", "normal"])
@pytest.mark.parametrize("u0", [0.0, -0.5])
def test_outflow_conditions_never_draw_water_in(channel, down, u0: float) -> None:  # noqa: ANN001
    g, ends = channel
    ref = _oracle(g, ends)
    h0 = np.full(g.N, 0.5, dtype=np.float32)
    ref.set_state(h0, (u0 * h0).astype(np.float32), np.zeros(g.N, np.float32))
    ref.set_open_conditions(*_conditions(g, ends, (WALL, 0.0), down))
    _run(ref, 0.0, 60.0)
    assert ref.open_volumes()[1, 1] == 0.0


def test_an_outlet_that_would_draw_water_in_is_a_wall(channel) -> None:  # noqa: ANN001
    g, ends = channel
    out = []
    for down in ((WALL, 0.0), (NORMAL, 1e-5)):
        ref = _oracle(g, ends)
        h0 = np.full(g.N, 0.5, dtype=np.float32)
        ref.set_state(h0, (-2.0 * h0).astype(np.float32), np.zeros(g.N, np.float32))
        ref.set_open_conditions(*_conditions(g, ends, (WALL, 0.0), down))
        _run(ref, 0.0, 1.0)
        out.append([ref.download(f, g.N) for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV)])
        assert not ref.open_volumes().any()
    for a, b in zip(*out, strict=True):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_a_dry_outflow_passes_nothing(channel) -> None:  # noqa: ANN001
    g, ends = channel
    ref = _oracle(g, ends)
    z = np.zeros(g.N, dtype=np.float32)
    ref.set_state(z, z, z)
    ref.set_open_conditions(*_conditions(g, ends, (WALL, 0.0), (FREE, 0.0)))
    _run(ref, 0.0, 60.0)
    assert not ref.open_volumes().any() and not ref.download(L.FIELD_H, g.N).any()


@pytest.mark.parametrize("up", [(INFLOW, 50.0), (STAGE, 10.4)], ids=["inflow", "stage"])
def test_water_entering_a_dry_channel_is_the_same_with_skipping_off(channel, up) -> None:  # noqa: ANN001
    g, ends = channel
    res = []
    for flags in (FLAGS_DEFAULT, FLAGS_DEFAULT & ~L.FLAG_ACTIVE_SET):
        ref = _oracle(g, ends, flags)
        z = np.zeros(g.N, dtype=np.float32)
        ref.set_state(z, z, z)
        ref.set_open_conditions(*_conditions(g, ends, up, (NORMAL, SLOPE)))
        steps = _run(ref, 0.0, 60.0)
        if flags & L.FLAG_ACTIVE_SET:
            assert (ref.download(L.FIELD_BLOCK_ACTIVE, ref.NB, np.uint32) == 0).any()
        res.append((steps, ref.download(L.FIELD_H, g.N).tobytes(), ref.open_volumes().tobytes()))
    assert res[0] == res[1]
    assert np.frombuffer(res[0][1], np.float32).any()


def test_a_stage_is_seen_by_the_first_step(channel) -> None:  # noqa: ANN001
    g, ends = channel
    ref = _oracle(g, ends)
    h0 = np.full(g.N, 0.01, dtype=np.float32)
    z = np.zeros(g.N, dtype=np.float32)
    ref.set_state(h0, z, z)
    ref.set_open_conditions(*_conditions(g, ends, (STAGE, 30.0), (WALL, 0.0)))
    assert _run(ref, 0.0, 1.0) >= 3


def test_inflow_forces_blocks_only_while_it_flows(channel) -> None:  # noqa: ANN001
    g, ends = channel
    ref = _oracle(g, ends)
    kinds = np.full(ref.n_open, INFLOW, dtype=np.uint32)
    ref.set_open_conditions(kinds, np.zeros(ref.n_open, np.float32))
    assert not ref.download(L.FIELD_BLOCK_FORCED, ref.NB, np.uint32).any()
    ref.set_open_conditions(*_conditions(g, ends, (INFLOW, 5.0), (WALL, 0.0)))
    forced = ref.download(L.FIELD_BLOCK_FORCED, ref.NB, np.uint32)
    assert forced.any()
    assert set(np.flatnonzero(forced)) == set(g.edge_cellL[ends[0]] // 256)
