"""Batched scenarios (needs an MI355X).

Several scenarios on meshes of different sizes advance together, each with its
own state, sources, clock and CFL dt. Each scenario's result must be bit for bit
the one it gets alone, and the FP32 oracle's. Stop times differ, so some
scenarios sit at their stop (no-op steps, frozen active set) while others move
on, and a retarget moves them again.
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays
from inundation_amd.mesh.geometry import build_geometry, hilbert_reorder
from inundation_amd.mesh.synth import lattice_tri_quad
from inundation_amd.tuning import dt_to_ticks
from oracle import cpu_ref
from tests import _drive as D
from tests._meshes import CASES, DRY_TOL, f32_state, valley_grid
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu


def _lattice(nq: int, seed: int):  # noqa: ANN202
    rng = np.random.default_rng(seed)
    v, f, o = lattice_tri_quad(nq, 4.0 * nq, rng)
    zb = np.zeros(o.size - 1, np.float32)
    for c in range(o.size - 1):
        pts = v[f[o[c]:o[c + 1]]]
        zb[c] = np.float32(0.3 * np.sin(pts[:, 0].mean() / 7.0) - 0.001 * pts[:, 0].mean())
    return hilbert_reorder(build_geometry(v, f, o, zb_from_file=zb), verbose=False)[0]


@pytest.fixture(scope="module")
def meshes():  # noqa: ANN201
    return [valley_grid(128, 64, 2.0), valley_grid(40, 24, 3.0), _lattice(24, 5)]


SCEN_MESH = [0, 1, 0, 2, 1]


def _scenario_inputs(geoms, k: int):  # noqa: ANN001, ANN202
    g = geoms[SCEN_MESH[k]]
    case = ["dam_break", "draining", "wet_lake", "dam_break", "dry_momentum"][k]
    h, hu, hv = f32_state(*CASES[case](g))
    n = np.full(g.N, 0.02 + 0.01 * k, np.float32)
    rng = np.random.default_rng(100 + k)
    left = g.centroid[:, 0] < 0.25 * float(g.cell_vertices[:, 0].max())
    cells = np.flatnonzero((rng.random(g.N) < 0.3) & left).astype(np.uint32)
    rates = (rng.uniform(0.0, 2e-3, cells.size) if k % 2 == 0 else np.full(cells.size, 5e-4)).astype(np.float32)
    stop = [4.5, 12.0, 7.5, 9.0, 2.25][k]
    return g, h, hu, hv, n, cells, rates, stop


def _ctrl(stop: float, dt_max: float, dt_init: float) -> L.InuCtrl:
    c = L.InuCtrl()
    c.now_ticks = 0
    c.stop_ticks = round(stop * 2**20)
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(dt_max)
    c.dry_dt_ticks = dt_to_ticks(min(dt_init, dt_max))
    c.cfl_safety = 1.0
    return c


def _run_alone(be, g, h, hu, hv, n, cells, rates, stop, blocks, extra):  # noqa: ANN001, ANN202
    D.setup(be, g.N, h, hu, hv, n, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates)
    be.ctrl_write(_ctrl(stop, 0.5, 0.05))
    be.prologue()
    out = []
    for nsteps in blocks:
        be.steps(nsteps)
        out.append(D.snapshot(be, g.N, -(-g.N // 256)))
    c = be.ctrl_read()
    c.stop_ticks = int(c.stop_ticks) + round(extra * 2**20)
    be.ctrl_write(c)
    be.retarget()
    be.steps(40)
    out.append(D.snapshot(be, g.N, -(-g.N // 256)))
    return out


def _batch_snapshot(ctx, k: int) -> dict[str, np.ndarray]:
    info = ctx.scenario_info(k)
    n_all = sum(ctx.scenario_info(s)["cells"] for s in range(ctx.n_scenarios))
    nb_all = sum(ctx.scenario_info(s)["blocks"] for s in range(ctx.n_scenarios))
    c0, n, b0, nb = info["cell0"], info["cells"], info["block0"], info["blocks"]
    out = {
        name: ctx.download(f, n_all)[c0:c0 + n]
        for name, f in (("h", L.FIELD_H), ("hu", L.FIELD_HU), ("hv", L.FIELD_HV), ("h_lo", L.FIELD_H_LO))
    }
    out["wet"] = ctx.download(L.FIELD_BLOCK_WET, nb_all, np.uint32)[b0:b0 + nb]
    out["active"] = ctx.download(L.FIELD_BLOCK_ACTIVE, nb_all, np.uint32)[b0:b0 + nb]
    c = ctx.ctrls_read()[k]
    out["ctrl"] = np.array([c.now_ticks, c.steps, c.dt_ticks], dtype=np.uint64)
    out["dt"] = ctx.dt_words()[k, :2].copy()
    return out


def test_batched_scenarios_are_bitwise_each_alone(meshes) -> None:  # noqa: ANN001
    blocks = [30, 50, 60]
    extra = 1.0
    ins = [_scenario_inputs(meshes, k) for k in range(len(SCEN_MESH))]
    dev = [device_mesh_arrays(g) for g in meshes]

    ctx = gpu_context()
    try:
        ctx.set_batch(dev, SCEN_MESH)
        assert ctx.n_scenarios == len(SCEN_MESH)
        cat = lambda j: np.concatenate([x[j] for x in ins])  # noqa: E731
        base = np.cumsum([0] + [x[0].N for x in ins])
        cells = np.concatenate([x[5] + np.uint32(base[k]) for k, x in enumerate(ins)]).astype(np.uint32)
        D.setup(ctx, int(base[-1]), cat(1), cat(2), cat(3), cat(4), dry_tol=DRY_TOL, src_cells=cells,
                src_rates=cat(6))
        ctx.ctrl_write([_ctrl(x[7], 0.5, 0.05) for x in ins])
        ctx.prologue()
        got = []
        for nsteps in blocks:
            ctx.steps(nsteps)
            got.append([_batch_snapshot(ctx, k) for k in range(len(ins))])
        cs = ctx.ctrls_read()
        for c in cs:
            c.stop_ticks = int(c.stop_ticks) + round(extra * 2**20)
        ctx.ctrl_write(cs)
        ctx.retarget()
        ctx.steps(40)
        got.append([_batch_snapshot(ctx, k) for k in range(len(ins))])
    finally:
        ctx.close()

    alone = gpu_context()
    try:
        for k, x in enumerate(ins):
            m = dev[SCEN_MESH[k]]
            alone.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"],
                           m["cidx"])
            want = _run_alone(alone, *x, blocks, extra)
            ref = _run_alone(cpu_ref.CpuRef(m), *x, blocks, extra)
            for j in range(len(want)):
                D.assert_same(got[j][k], want[j], f"scenario {k} snapshot {j} (batch vs alone)")
                D.assert_same(want[j], ref[j], f"scenario {k} snapshot {j} (alone vs oracle)")
    finally:
        alone.close()
    # The stops differ: some scenarios were frozen while others stepped.
    steps = [int(got[2][k]["ctrl"][1]) for k in range(len(ins))]
    assert len(set(steps)) > 1, steps


def test_source_rates_update_is_a_new_source_list(meshes) -> None:  # noqa: ANN001
    """inu_set_source_rates (rates of the same cells, one turned off) equals a new inu_set_sources."""
    g = meshes[0]
    m = device_mesh_arrays(g)
    h, hu, hv = f32_state(*CASES["dam_break"](g))
    left = g.centroid[:, 0] < 0.25 * float(g.cell_vertices[:, 0].max())
    cells = np.flatnonzero(left).astype(np.uint32)
    rates0 = np.full(cells.size, 1e-3, np.float32)
    rates1 = np.linspace(0.0, 3e-3, cells.size).astype(np.float32)  # the first cell's source is turned off
    nb = -(-g.N // 256)
    ctx = gpu_context()
    try:
        ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
        ref = cpu_ref.CpuRef(m)
        for be in (ctx, ref):
            D.setup(be, g.N, h, hu, hv, 0.03, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates0)
            be.ctrl_write(_ctrl(6.0, 0.5, 0.05))
            be.prologue()
            be.steps(30)
        D.assert_same(D.snapshot(ctx, g.N, nb), D.snapshot(ref, g.N, nb), "before the update")
        ctx.set_source_rates(cells, rates1)
        ref.set_sources(cells, rates1)
        for be in (ctx, ref):
            be.steps(40)
        D.assert_same(D.snapshot(ctx, g.N, nb), D.snapshot(ref, g.N, nb), "after the update")
        with pytest.raises(ValueError, match="not a source cell"):
            ctx.set_source_rates(np.array([int(np.flatnonzero(~left)[0])], np.uint32), np.ones(1, np.float32))
    finally:
        ctx.close()
