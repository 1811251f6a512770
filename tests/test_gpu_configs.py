"""The benchmark configurations themselves on the HIP path (needs an MI355X).

- C1, lake at rest (4096 quads over 12 bumps, 1000 Heun steps at CFL 0.5):
  bitwise the FP32 oracle over the whole run; max(|u|, |v|) over cells deeper
  than 10 dry_tol below 1e-6 m/s (the metric of the reference's
  ``bench/lake_at_rest.py:197-212``); the FP64 restatement replayed with the
  same dt sequence within the stated tolerance.
- C3, the 16 003 008-cell rainfall flood: the wetting transient (the first 20
  steps from dry) and a wet window (10 steps from the state 100 steps in)
  bitwise the FP32 oracle; 4 steps of the FP64 restatement on a slice of the
  mesh (the core cells plus the 8 rings they depend on) within the stated
  tolerance.
- C2, the 1 M-cell Voronoi inflow flood: the first three one-minute
  hydrograph phases bitwise the FP32 oracle, phase by phase.
- C4, 64 scenarios over five Voronoi meshes of 50 k to 2 M cells: ten of
  them (two per mesh) batched through two phases at the autotuned tile
  level, each bit for bit the same scenario run alone and the FP32 oracle.
- C5, the 250 k-cell adjoint calibration mesh: the GPU gradient of a gauge
  misfit over the first 10 steps against the FP64 adjoint, and the replay of
  the recording bit for bit.

Each result's error against FP64 is printed with the ``[configs]`` lines.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays
from inundation_amd.tuning import dt_to_ticks, seconds_to_ticks
from oracle import cpu_ref, ref64
from tests import _drive as D
from tests._slices import ring_closure, submesh
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu

FP64_DEPTH_TOL = 1e-5  # m, after the replayed steps
FP64_MOMENTUM_TOL = 1e-4  # m^2/s


@pytest.fixture(scope="module")
def say(request):  # noqa: ANN001, ANN201
    """Print past pytest's capture, so long steps show progress."""
    capman = request.config.pluginmanager.getplugin("capturemanager")

    def f(msg: str) -> None:
        with capman.global_and_fixture_disabled():
            print(f"[configs] {msg}", flush=True)

    return f


@pytest.fixture(scope="module", autouse=True)
def oracle_threads() -> None:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cores = os.cpu_count() or 1
    cpu_ref.set_threads(max(1, min(cores, 16)))


def _ctrl(now: int, stop: int, dt_max: float, dt_init: float) -> L.InuCtrl:
    c = L.InuCtrl()
    c.now_ticks = now
    c.stop_ticks = stop
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(dt_max)
    c.dry_dt_ticks = dt_to_ticks(min(dt_init, dt_max))
    c.cfl_safety = 1.0
    return c


def _set_mesh(ctx, m) -> None:  # noqa: ANN001
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])


def _fp64_errors(got: dict, want: ref64.State64, cells=None) -> tuple[float, float]:  # noqa: ANN001
    sel = slice(None) if cells is None else cells
    h = got["h"][sel].astype(np.float64) + got["h_lo"][sel].astype(np.float64)
    eh = float(np.abs(h - want.h).max())
    eq = float(max(np.abs(got["hu"][sel] - want.hu).max(), np.abs(got["hv"][sel] - want.hv).max()))
    return eh, eq


# -- C1 ---------------------------------------------------------------------------


def test_c1_lake_at_rest(say) -> None:  # noqa: ANN001
    from inundation_amd.workloads import lake_at_rest

    wl = lake_at_rest()
    n_cells, nb = wl.geom.N, -(-wl.geom.N // 256)
    m = device_mesh_arrays(wl.geom)
    ctx = gpu_context()
    try:
        _set_mesh(ctx, m)
        ref = cpu_ref.CpuRef(m)
        for be in (ctx, ref):
            D.setup(be, n_cells, wl.h0, wl.hu0, wl.hv0, wl.manning, dry_tol=wl.dry_tol, cfl=wl.cfl)
            be.ctrl_write(_ctrl(0, 2**40, wl.dt_max, wl.dt_init))
            be.prologue()
        dts = []
        for _ in range(wl.steps):
            dts.append(int(ref.ctrl_read().dt_ticks) / 2**20)
            ref.steps(1)
        for _ in range(wl.steps // 200):
            ctx.steps(200, use_graph=True)
        got = D.snapshot(ctx, n_cells, nb)
        D.assert_same(got, D.snapshot(ref, n_cells, nb), "C1 after 1000 steps")
        assert int(got["ctrl"][1]) == wl.steps
        h = got["h"].astype(np.float64)
        wet = h > 10 * wl.dry_tol
        umax = float(max((np.abs(got["hu"][wet]) / h[wet]).max(), (np.abs(got["hv"][wet]) / h[wet]).max()))
        m64 = ref64.Mesh64.from_device_arrays(m)
        s0 = ref64.State64(wl.h0.astype(np.float64), np.zeros(n_cells), np.zeros(n_cells))
        want = ref64.run_dt_sequence(m64, s0, wl.manning.astype(np.float64), dts, 9.81, wl.dry_tol)
        eh, eq = _fp64_errors(got, want)
        say(f"C1 lake at rest: 1000 steps ({sum(dts):.1f} s) bitwise the FP32 oracle; max |u| {umax:.2e} m/s over "
            f"{int(wet.sum())} wet cells; vs FP64: depth {eh:.2e} m, momentum {eq:.2e} m^2/s")
        assert umax < 1e-6
        assert eh <= FP64_DEPTH_TOL and eq <= FP64_MOMENTUM_TOL
    finally:
        ctx.close()


# -- C3 ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c3(say):  # noqa: ANN001, ANN201
    from inundation_amd.workloads import rainfall

    say("C3: building the 16 003 008-cell rainfall mesh")
    wl = rainfall()
    m = device_mesh_arrays(wl.geom)
    ctx = gpu_context()
    _set_mesh(ctx, m)
    say("C3: mesh on the device; building the oracle")
    ref = cpu_ref.CpuRef(m)
    cells, rates = wl.source_pairs()
    for be in (ctx, ref):
        D.setup(be, wl.geom.N, wl.h0, wl.hu0, wl.hv0, wl.manning, dry_tol=wl.dry_tol, cfl=wl.cfl,
                src_cells=cells, src_rates=rates)
    yield wl, m, ctx, ref
    ctx.close()


def test_c3_wetting_transient_is_bitwise_the_oracle(c3, say) -> None:  # noqa: ANN001
    wl, m, ctx, ref = c3
    n_cells, nb = wl.geom.N, -(-wl.geom.N // 256)
    for be in (ctx, ref):
        be.set_state(wl.h0, wl.hu0, wl.hv0, None)
        be.ctrl_write(_ctrl(0, seconds_to_ticks(1800.0), wl.dt_max, wl.dt_init))
        be.prologue()
    for k in range(4):
        ctx.steps(5, use_graph=True)
        ref.steps(5)
        got = D.snapshot(ctx, n_cells, nb)
        D.assert_same(got, D.snapshot(ref, n_cells, nb), f"C3 from dry, after {5 * (k + 1)} steps")
        say(f"C3 from dry: {5 * (k + 1)} steps bitwise the oracle, {int(got['active'].sum())} of {nb} blocks active")


def test_c3_wet_window_is_bitwise_the_oracle_and_near_fp64(c3, say) -> None:  # noqa: ANN001
    wl, m, ctx, ref = c3
    n_cells, nb = wl.geom.N, -(-wl.geom.N // 256)
    ctx.set_state(wl.h0, wl.hu0, wl.hv0, None)
    ctx.ctrl_write(_ctrl(0, seconds_to_ticks(1800.0), wl.dt_max, wl.dt_init))
    ctx.prologue()
    ctx.steps(100, use_graph=True)
    state = tuple(ctx.download(f, n_cells) for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO))
    now = int(ctx.ctrl_read().now_ticks)
    for be in (ctx, ref):
        be.set_state(*state)
        be.ctrl_write(_ctrl(now, now + seconds_to_ticks(600.0), wl.dt_max, wl.dt_init))
        be.prologue()
    # FP64 slice: the cells of 16 blocks in the middle of the mesh and the 8
    # rings that 4 steps depend on.
    core = np.arange(n_cells // 2, n_cells // 2 + 16 * 256, dtype=np.int64)
    sm, cells = submesh(m, ring_closure(m, core, 8))
    dts = []
    for _ in range(4):
        dts.append(int(ctx.ctrl_read().dt_ticks) / 2**20)
        ctx.steps(1, use_graph=False)
    got4 = D.snapshot(ctx, n_cells, nb)
    ctx.steps(6, use_graph=True)
    ref.steps(10)
    got = D.snapshot(ctx, n_cells, nb)
    D.assert_same(got, D.snapshot(ref, n_cells, nb), "C3 wet window, 10 steps")
    assert got["active"].all()
    say(f"C3 wet window: 10 steps from step 100 (t = {now / 2**20:.1f} s) bitwise the oracle")
    # FP64 replay of the first 4 of those steps on the slice, from the same state.
    src = np.zeros(n_cells)
    cells_src, rates = wl.source_pairs()
    src[cells_src] = rates.astype(np.float64)
    s0 = ref64.State64(state[0][cells].astype(np.float64) + state[3][cells].astype(np.float64),
                       state[1][cells].astype(np.float64), state[2][cells].astype(np.float64))
    want = ref64.run_dt_sequence(ref64.Mesh64.from_device_arrays(sm), s0, wl.manning[cells].astype(np.float64), dts,
                                 9.81, wl.dry_tol, src=src[cells])
    pos = np.searchsorted(cells, core)
    sub = ref64.State64(want.h[pos], want.hu[pos], want.hv[pos])
    eh, eq = _fp64_errors(got4, sub, core)
    say(f"C3 vs FP64 over 4 steps ({sum(dts):.2f} s) on {core.size} cells ({cells.size} with their rings): depth "
        f"{eh:.2e} m, momentum {eq:.2e} m^2/s")
    assert eh <= FP64_DEPTH_TOL and eq <= FP64_MOMENTUM_TOL


# -- C2 ---------------------------------------------------------------------------


def test_c2_first_inflow_phases_are_bitwise_the_oracle(say) -> None:  # noqa: ANN001
    from inundation_amd.workflow import build_phase_source_rate
    from inundation_amd.workloads import inflow

    say("C2: building the 1 M-cell Voronoi mesh")
    wl, phases = inflow()
    geom = wl.geom
    n_cells, nb = geom.N, -(-geom.N // 256)
    m = device_mesh_arrays(geom)
    ctx = gpu_context()
    try:
        _set_mesh(ctx, m)
        ref = cpu_ref.CpuRef(m)
        for be in (ctx, ref):
            D.setup(be, n_cells, wl.h0, wl.hu0, wl.hv0, wl.manning, dry_tol=wl.dry_tol, cfl=wl.cfl)
        valid = geom.area > 1e-6
        stop = 0
        for k, ph in enumerate(phases[:3]):
            rate = build_phase_source_rate(geom, ph, valid)
            dh = np.asarray(rate, np.float32) * (1.0 / geom.area).astype(np.float32)
            cells = np.flatnonzero(dh).astype(np.uint32)
            stop += seconds_to_ticks(ph.duration_s)
            for be in (ctx, ref):
                be.set_sources(cells, dh[cells].astype(np.float32))
                c = be.ctrl_read() if k else _ctrl(0, 0, wl.dt_max, wl.dt_init)
                c.stop_ticks = stop
                be.ctrl_write(c)
                be.retarget() if k else be.prologue()
            while int(ctx.ctrl_read().now_ticks) < stop:
                ctx.steps(50, use_graph=True)
                ref.steps(50)
            got = D.snapshot(ctx, n_cells, nb)
            D.assert_same(got, D.snapshot(ref, n_cells, nb), f"C2 after phase {k + 1}")
            say(f"C2 phase {k + 1}: {int(got['ctrl'][1])} steps to t = {int(got['ctrl'][0]) / 2**20:.0f} s bitwise "
                f"the oracle; {int((got['h'] >= wl.dry_tol).sum())} wet cells")
    finally:
        ctx.close()


# -- C5 ---------------------------------------------------------------------------


def test_c5_calibration_gradient_matches_fp64_and_replays_bitwise(say) -> None:  # noqa: ANN001
    from inundation_amd.hip.adjoint import GaugeRun
    from inundation_amd.workloads import calibration
    from oracle import adjoint64 as A

    say("C5: building the 250 k-cell calibration mesh")
    cal = calibration()
    geom = cal.geom
    m = device_mesh_arrays(geom)
    ctx = gpu_context()
    try:
        _set_mesh(ctx, m)
        k_steps, every = 10, 5
        d2 = (geom.centroid[:, 0] - geom.centroid[cal.inflow_cell, 0]) ** 2 + (
            geom.centroid[:, 1] - geom.centroid[cal.inflow_cell, 1]
        ) ** 2
        gauges = np.sort(np.argsort(d2)[:32]).astype(np.uint32)
        zero = np.zeros(geom.N, np.float32)
        run = GaugeRun(ctx=ctx, state0=(zero, zero, zero), gauges=gauges, every=every, n_steps=k_steps, seg=5,
                       source_cells=cal.source_cells(), rates_at=cal.rates_at, graph_steps=k_steps)
        n0 = np.full(geom.N, cal.n_start, np.float32)
        z = geom.zb.astype(np.float32)
        series = run.forward(n0, z)
        gn, gz = run.backward((2.0 * series.astype(np.float64)).astype(np.float32))
        assert np.array_equal(run.replay(n0, z).view(np.uint32), series.view(np.uint32)), "replay differs"
        dts = run.dts().astype(np.float64)
        src = np.zeros(geom.N)
        src[cal.inflow_cell] = float(cal.rates_at(0.0)[0])
        _, gz64, gn64 = A.misfit_and_grads(m, z.astype(np.float64), n0.astype(np.float64),
                                           tuple(np.zeros(geom.N) for _ in range(3)), dts, src, gauges,
                                           np.zeros((k_steps // every, gauges.size)), every, thin=100.0 * run.dry_tol)
        en = float(np.linalg.norm(gn - gn64) / np.linalg.norm(gn64))
        ez = float(np.linalg.norm(gz - gz64) / np.linalg.norm(gz64))
        say(f"C5: gradient of {k_steps} recorded steps vs the FP64 adjoint: n {en:.2e}, z {ez:.2e} (relative norm); "
            f"replay bitwise")
        assert en < 1e-4 and ez < 1e-4
    finally:
        ctx.close()


# -- C4 ---------------------------------------------------------------------------


def _drive_phases(be, parts, n_phases: int, multi: bool) -> None:  # noqa: ANN001
    """One-minute phases on a context or oracle holding the scenarios ``parts`` back to back.

    ``parts`` is [(scenario, geometry)]; steps go in blocks of 50. A
    scenario that reaches the end of its phase moves on at once, as the
    bench's batched run does: every scenario's sources for its own phase's
    mid-time, its stop one minute on, and a retarget, which leaves the
    scenarios still inside a phase where they were. Alone, this is the
    plain phase loop.
    """
    base = np.cumsum([0] + [g.N for _, g in parts])
    phase = [0] * len(parts)

    def set_sources() -> None:
        pairs = [s.source_pairs(g, (phase[k] + 0.5) * 60.0) for k, (s, g) in enumerate(parts)]
        be.set_sources(np.concatenate([c + np.uint32(base[k]) for k, (c, _) in enumerate(pairs)]).astype(np.uint32),
                       np.concatenate([r for _, r in pairs]).astype(np.float32))

    def ctrls() -> list:  # noqa: ANN202
        return be.ctrls_read() if multi else [be.ctrl_read()]

    set_sources()
    cs = ctrls()
    for c in cs:
        c.stop_ticks = 60 * 2**20
    be.ctrl_write(cs if multi else cs[0])
    be.prologue()
    while True:
        be.steps(50, use_graph=True)
        cs = ctrls()
        at_stop = [int(c.now_ticks) >= int(c.stop_ticks) for c in cs]
        move = [k for k in range(len(parts)) if at_stop[k] and phase[k] + 1 < n_phases]
        if not move:
            if all(at_stop):
                return
            continue
        for k in move:
            phase[k] += 1
            cs[k].stop_ticks = (phase[k] + 1) * 60 * 2**20
        set_sources()
        be.ctrl_write(cs if multi else cs[0])
        be.retarget()


def test_c4_batched_scenarios_on_the_five_meshes_are_each_alone_and_the_oracle(say) -> None:  # noqa: ANN001
    from inundation_amd.workloads import batched

    say("C4: building the five Voronoi meshes (50 k to 2 M cells)")
    geoms, scen = batched(n_scenarios=64, progress=say, processes=8)
    picked = list(range(10))  # two scenarios on each mesh
    parts = [(scen[k], geoms[scen[k].mesh]) for k in picked]
    dev = [device_mesh_arrays(g) for g in geoms]
    n_phases = 2
    zero_ctrl = _ctrl(0, 0, 1.0, 0.1)
    ctx = gpu_context()
    try:
        ctx.set_batch(dev, [s.mesh for s, _ in parts])
        n_all = sum(g.N for _, g in parts)
        D.setup(ctx, n_all, *(np.zeros(n_all, np.float32) for _ in range(3)), 0.035, dry_tol=1e-4, cfl=0.8)
        ctx.ctrl_write([_ctrl(0, 0, 1.0, 0.1) for _ in parts])
        # Autotune on a flooding state, then the batched run from the start at that tile level.
        pairs = [s.source_pairs(g, 30.0) for s, g in parts]
        base = np.cumsum([0] + [g.N for _, g in parts])
        ctx.set_sources(np.concatenate([c + np.uint32(base[k]) for k, (c, _) in enumerate(pairs)]).astype(np.uint32),
                        np.concatenate([r for _, r in pairs]).astype(np.float32))
        cs = ctx.ctrls_read()
        for c in cs:
            c.stop_ticks = 10**6 * 2**20
        ctx.ctrl_write(cs)
        ctx.prologue()
        ctx.steps(100, use_graph=False)
        tile = ctx.autotune(3)["tile"]
        z = np.zeros(n_all, np.float32)
        ctx.set_state(z, z, z, None)
        ctx.ctrl_write([_ctrl(0, 0, 1.0, 0.1) for _ in parts])
        _drive_phases(ctx, parts, n_phases, multi=True)
        say(f"C4: {len(parts)} scenarios batched over {n_phases} phases at tile level {tile}: "
            f"{[int(c.steps) for c in ctx.ctrls_read()]} steps")
        fields = (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO)
        got = {f: ctx.download(f, n_all) for f in fields}
        ctrls = ctx.ctrls_read()
        for k, (s, g) in enumerate(parts):
            alone = gpu_context()
            try:
                _set_mesh(alone, dev[s.mesh])
                ref = cpu_ref.CpuRef(dev[s.mesh])
                for be in (alone, ref):
                    D.setup(be, g.N, *(np.zeros(g.N, np.float32) for _ in range(3)), 0.035, dry_tol=1e-4, cfl=0.8)
                    be.ctrl_write(zero_ctrl)
                alone.set_tile(tile)
                _drive_phases(alone, [(s, g)], n_phases, multi=False)
                _drive_phases(ref, [(s, g)], n_phases, multi=False)
                for f in fields:
                    mine = got[f][base[k]:base[k + 1]]
                    for other, who in ((alone, "alone"), (ref, "oracle")):
                        o = other.download(f, g.N)
                        d = np.flatnonzero(mine.view(np.uint32) != o.view(np.uint32))
                        assert d.size == 0, f"scenario {picked[k]} field {f}: {d.size} words differ from {who}"
                for other, who in ((alone, "alone"), (ref, "oracle")):
                    assert bytes(other.ctrl_read()) == bytes(ctrls[k]), f"scenario {picked[k]}: clock differs from {who}"
            finally:
                alone.close()
            say(f"C4 scenario {picked[k]} ({g.N} cells, {int(ctrls[k].steps)} steps): batched = alone = oracle")
    finally:
        ctx.close()
