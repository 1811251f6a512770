"""The HIP adjoint (needs an MI355X).

- the edge-word VJP against FP64 reverse-mode autodiff of the scheme's edge
  words (``oracle/adjoint64.py``) on random face states: dry sides, the
  dry-bed blend band, stepped faces and walls;
- whole-run gradients of a gauge misfit with respect to Manning n and bed z
  against the FP64 reference gradient of the same run (same recorded dt
  sequence), to 1e-3 relative in norm, over several checkpoint segments;
- the replay of the recording reproduces the recorded gauge series bit for
  bit, and GPU central differences of the replayed run agree with the adjoint.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from inundation_amd.hip import lib as L
from inundation_amd.hip.adjoint import GaugeRun, gauge_depths
from inundation_amd.hip.solver import device_mesh_arrays
from oracle import adjoint64 as A
from tests._meshes import valley_grid
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu
G = 9.81
THIN = 100.0 * 1e-4  # the reverse pass's default thin-layer depth at dry_tol 1e-4


@pytest.fixture(scope="module")
def ctx():  # noqa: ANN201
    c = gpu_context()
    yield c
    c.close()


def _edge_cases(n: int, dry: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, 17), np.float32)
    kind = rng.integers(0, 4, size=(n, 2))
    depth = np.where(kind == 0, 0.0, np.where(kind == 1, rng.uniform(0.2, 0.9, (n, 2)) * dry,
                     np.where(kind == 2, rng.uniform(1.05, 1.95, (n, 2)) * dry, rng.uniform(0.01, 2.0, (n, 2)))))
    vel = rng.uniform(-2.0, 2.0, (n, 4))
    ang = rng.uniform(0, 2 * np.pi, n)
    wall = rng.random(n) < 0.15
    step = rng.uniform(0.0, 0.3, n) * (rng.random(n) < 0.7)
    left_step = rng.random(n) < 0.5
    rows[:, 0] = depth[:, 0]
    rows[:, 1] = vel[:, 0] * depth[:, 0]
    rows[:, 2] = vel[:, 1] * depth[:, 0]
    rows[:, 3] = depth[:, 1]
    rows[:, 4] = vel[:, 2] * depth[:, 1]
    rows[:, 5] = vel[:, 3] * depth[:, 1]
    rows[:, 6] = np.where(left_step & ~wall, step, 0.0)
    rows[:, 7] = np.where(~left_step & ~wall, step, 0.0)
    rows[:, 8] = np.cos(ang)
    rows[:, 9] = np.sin(ang)
    rows[:, 10] = rng.uniform(0.5, 3.0, n)
    rows[:, 11] = wall
    rows[:, 12:17] = rng.normal(0.0, 1.0, (n, 5))
    return rows


def _edge_vjp64(rows: np.ndarray, dry: float, thin: float) -> np.ndarray:
    n = rows.shape[0]
    r = rows.astype(np.float64)
    wall = torch.as_tensor(rows[:, 11] != 0)
    cl = torch.arange(0, 2 * n, 2)
    cr = torch.where(wall, torch.full_like(cl, -1), cl + 1)
    m = A.MeshT(cl=cl, cr=cr, crs=torch.where(wall, cl, cl + 1), wall=wall, nx=torch.as_tensor(r[:, 8]),
                ny=torch.as_tensor(r[:, 9]), length=torch.as_tensor(r[:, 10]), area=torch.ones(2 * n, dtype=A.F64),
                slot_cell=torch.zeros(0, dtype=torch.long), slot_edge=torch.zeros(0, dtype=torch.long),
                slot_right=torch.zeros(0, dtype=torch.bool))
    h = torch.tensor(np.stack([r[:, 0], r[:, 3]], 1).reshape(-1), requires_grad=True)
    hu = torch.tensor(np.stack([r[:, 1], r[:, 4]], 1).reshape(-1), requires_grad=True)
    hv = torch.tensor(np.stack([r[:, 2], r[:, 5]], 1).reshape(-1), requires_grad=True)
    dzl = torch.tensor(r[:, 6], requires_grad=True)
    dzr = torch.tensor(r[:, 7], requires_grad=True)
    w = A.edge_words(m, h, hu, hv, dzl, dzr, G, dry, thin)
    (w * torch.as_tensor(r[:, 12:17].T)).sum().backward()
    gh, ghu, ghv = (x.grad.numpy().reshape(n, 2) for x in (h, hu, hv))
    return np.column_stack([gh[:, 0], ghu[:, 0], ghv[:, 0], gh[:, 1], ghu[:, 1], ghv[:, 1], dzl.grad.numpy(),
                            dzr.grad.numpy()])


@pytest.mark.parametrize("thin", [None, 0.0])
def test_edge_vjp_matches_fp64_autodiff(ctx, thin) -> None:  # noqa: ANN001
    """Default thin-layer depth (100 dry_tol, many rows thin) and none."""
    dry = 1e-3
    rows = _edge_cases(20000, dry, 3)
    ctx.adj_set_thin(-1.0 if thin is None else thin)
    try:
        got = ctx.probe_vjp(rows, G, dry)
    finally:
        ctx.adj_set_thin(-1.0)
    want = _edge_vjp64(rows, dry, 100.0 * dry if thin is None else thin)
    wall = rows[:, 11] != 0
    want[wall, 6:] = 0.0  # walls have no bed step
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-30
    err = np.abs(got - want) / scale
    bad = np.flatnonzero(err.max(axis=1) > 2e-3)
    # A handful of rows sit on a branch the FP32 and FP64 evaluations take differently.
    assert bad.size <= 10, (bad.size, rows[bad[:3]].tolist(), got[bad[:3]].tolist(), want[bad[:3]].tolist())


def _run_case(ctx, *, wet: bool, n_steps: int = 60, seg: int = 16):  # noqa: ANN001, ANN202
    g = valley_grid(48, 32, 2.0)
    m = device_mesh_arrays(g)
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    x = g.centroid[:, 0].astype(np.float64)
    y = g.centroid[:, 1].astype(np.float64)
    z = (g.zb.astype(np.float64) + 2e-3 * np.sin(3.7 * x + 1.3 * y)).astype(np.float32)
    if wet:
        h = np.maximum(float(z.max()) + 0.3 - z, 0.0)
        hu = 0.2 * np.sin(2 * np.pi * x / 96.0) * h
        hv = 0.05 * np.cos(2 * np.pi * y / 64.0) * h
    else:
        h = np.where(x < 24.0, 0.8, 0.0) + np.where(np.abs(y - 32.0) < 6.0, 0.05, 0.0)
        hu = np.zeros_like(h)
        hv = np.zeros_like(h)
    state0 = tuple(np.asarray(a, np.float32) for a in (h, hu, hv))
    rng = np.random.default_rng(9)
    cells = np.flatnonzero(rng.random(g.N) < 0.3).astype(np.uint32)
    rates = np.full(cells.size, 2e-4, np.float32)
    gauges = np.sort(rng.choice(np.flatnonzero(h > 0.05), 8, replace=False)).astype(np.uint32)
    n = (0.02 + 0.03 * rng.random(g.N)).astype(np.float32)
    run = GaugeRun(ctx=ctx, state0=state0, gauges=gauges, every=5, n_steps=n_steps, seg=seg, dry_tol=1e-4,
                   dt_max=0.2, dt_init=0.05, source_cells=cells, phases=[(0, rates)], graph_steps=10)
    src = np.zeros(g.N)
    src[cells] = rates
    return g, m, z, n, state0, gauges, run, src


def test_adjoint_matches_the_fp64_reference_gradient(ctx) -> None:  # noqa: ANN001
    g, m, z, n, state0, gauges, run, src = _run_case(ctx, wet=True)
    nt = torch.tensor(n, requires_grad=True)
    zt = torch.tensor(z, requires_grad=True)
    series = gauge_depths(nt, zt, run)
    obs = series.detach() - 0.01 * torch.linspace(0.5, 1.5, series.numel()).reshape(series.shape)
    ((series - obs) ** 2).sum().backward()
    dts = run.dts()
    assert (dts > 0).all() and dts.size == 60
    _, gz64, gn64 = A.misfit_and_grads(m, z.astype(np.float64), n.astype(np.float64), state0, dts.astype(np.float64),
                                       src, gauges, obs.numpy().astype(np.float64), 5, thin=THIN)
    for got, want, name in ((nt.grad.numpy(), gn64, "n"), (zt.grad.numpy(), gz64, "z")):
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"all-wet valley, GPU vs FP64 gradient of {name}: {rel:.2e}")
        assert rel < 1e-3, (name, rel)


def test_replay_reproduces_the_recording_and_central_differences(ctx) -> None:  # noqa: ANN001
    g, m, z, n, state0, gauges, run, src = _run_case(ctx, wet=True)
    series = run.forward(n, z)
    assert np.array_equal(run.replay(n, z).view(np.uint32), series.view(np.uint32))
    obs = series - 0.01
    gn, gz = run.backward(2.0 * (series - obs))
    # Central differences of the replayed run (same dt sequence) at the cells of largest gradient.
    def loss(nn, zz):  # noqa: ANN001, ANN202
        s = run.replay(nn, zz).astype(np.float64)
        return float(((s - obs) ** 2).sum())

    worst = 0.0
    for vec, grad, eps in ((n, gn, 2e-3), (z, gz, 2e-3)):
        for i in np.argsort(-np.abs(grad))[:4].tolist():
            up, dn = vec.copy(), vec.copy()
            up[i] += eps
            dn[i] -= eps
            fd = (loss(up, z) - loss(dn, z)) / (2 * eps) if vec is n else (loss(n, up) - loss(n, dn)) / (2 * eps)
            worst = max(worst, abs(grad[i] - fd) / abs(fd))
    print(f"all-wet valley, worst GPU adjoint vs central difference: {worst:.2e}")
    assert worst < 1e-2, worst


def test_adjoint_with_wet_dry_fronts_follows_the_fp64_reference(ctx) -> None:  # noqa: ANN001
    g, m, z, n, state0, gauges, run, src = _run_case(ctx, wet=False, n_steps=50, seg=12)
    series = run.forward(n, z)
    obs = series - 0.01
    gn, gz = run.backward(2.0 * (series - obs))
    dts = run.dts()
    _, gz64, gn64 = A.misfit_and_grads(m, z.astype(np.float64), n.astype(np.float64), state0, dts.astype(np.float64),
                                       src, gauges, obs.astype(np.float64), 5, thin=THIN)
    for got, want, name in ((gn, gn64, "n"), (gz, gz64, "z")):
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"wet/dry valley, GPU vs FP64 gradient of {name}: {rel:.2e}")
        assert rel < 1e-4, (name, rel)


def test_adjoint_matches_fp64_central_differences_all_wet(ctx) -> None:  # noqa: ANN001
    """The GPU gradient against FP64 central differences of the same run (the
    recorded dt sequence), at the cells of largest gradient, to 1e-3."""
    g, m, z, n, state0, gauges, run, src = _run_case(ctx, wet=True, n_steps=40, seg=16)
    series = run.forward(n, z)
    obs = series.astype(np.float64) - 0.01 * np.linspace(0.5, 1.5, series.size).reshape(series.shape)
    gn, gz = run.backward((2.0 * (series.astype(np.float64) - obs)).astype(np.float32))
    dts = run.dts().astype(np.float64)
    z64, n64 = z.astype(np.float64), n.astype(np.float64)

    def loss(nn, zz):  # noqa: ANN001, ANN202
        s = A.gauge_series(m, zz, nn, state0, dts, src, gauges, 5)
        return float(((s - obs) ** 2).sum())

    worst = 0.0
    for vec, grad, eps in ((n64, gn, 1e-5), (z64, gz, 1e-6)):
        for i in np.argsort(-np.abs(grad))[:4].tolist():
            up, dn = vec.copy(), vec.copy()
            up[i] += eps
            dn[i] -= eps
            fd = (loss(up, z64) - loss(dn, z64)) / (2 * eps) if vec is n64 else (loss(n64, up) - loss(n64, dn)) / (2 * eps)
            worst = max(worst, abs(grad[i] - fd) / abs(fd))
    print(f"all-wet valley, worst GPU adjoint vs FP64 central difference: {worst:.2e}")
    assert worst < 1e-3, worst


@pytest.mark.parametrize("slots", [1, 2, 3, 8])
def test_binomial_checkpointing_gives_the_uniform_gradient_bitwise(ctx, slots: int) -> None:  # noqa: ANN001
    """Binomial checkpointing replays the same steps from checkpoints it stores itself: the same bits.

    Wet/dry valley, 60 steps in segments of 8 (the last one partial), source
    rates changing at step 16 (a segment boundary) and step 23 (inside a
    segment). The backward replays every step once with tape and advances the
    schedule's count of segments without it.
    """
    g, m, z, n, state0, gauges, run, src = _run_case(ctx, wet=False, n_steps=60, seg=8)
    rates = run.phases[0][1]
    run.phases = [(0, rates), (16, 2.0 * rates), (23, 0.5 * rates)]
    run.graph_steps = 8
    series = run.forward(n, z)
    seeds = (2.0 * (series - 0.9 * series)).astype(np.float32)
    gn_u, gz_u = run.backward(seeds)
    assert ctx.adj_replayed() == {"advanced": 0, "taped": 60}
    run.slots = slots
    try:
        series_b = run.forward(n, z)
        assert np.array_equal(series_b.view(np.uint32), series.view(np.uint32))
        gn_b, gz_b = run.backward(seeds)
    finally:
        run.slots = 0
    assert np.array_equal(gn_b.view(np.uint32), gn_u.view(np.uint32))
    assert np.array_equal(gz_b.view(np.uint32), gz_u.view(np.uint32))
    assert np.abs(gn_u).max() > 0.0 and np.abs(gz_u).max() > 0.0
    # 8 segments; every segment advanced without tape is a whole one.
    assert ctx.adj_replayed() == {"advanced": 8 * L.schedule_cost(8, slots), "taped": 60}


def test_a_new_mesh_after_a_recording_keeps_its_buffers(ctx) -> None:  # noqa: ANN001
    """Setting a mesh frees the recording with everything else; ending the recording later frees nothing live.

    The new mesh's allocations may reuse the old recording's addresses. Three
    rounds of record, differentiate, set the mesh again and end the
    recording; each following run must give the first run's gauge series
    bit for bit.
    """
    first = None
    for _ in range(3):
        g, m, z, n, state0, gauges, run, src = _run_case(ctx, wet=True, n_steps=20, seg=8)
        series = run.forward(n, z)
        run.backward(np.ones_like(series))
        if first is None:
            first = series
        assert np.array_equal(series.view(np.uint32), first.view(np.uint32))
        _set = device_mesh_arrays(g)
        ctx.set_mesh(_set["cell_l"], _set["cell_r"], _set["nx"], _set["ny"], _set["len"], _set["dz2"],
                     _set["area"], _set["cptr"], _set["cidx"])
        ctx.adj_end()
