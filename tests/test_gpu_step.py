"""HIP backend against the oracles (needs an MI355X).

- bitwise against the FP32 oracle (``oracle/cpu_ref.cpp``): state, low words,
  block flags, clock and dt words after fixed-dt and CFL-driven runs, with and
  without the active set, compensation and sources;
- within the reference's FP32 tolerances of the FP64 oracle
  (``tests/unit/test_gpu_step_parity.py:31-35, 161-163``): 2e-6 m depth and
  2e-5 momentum after 8 steps of dt = 2^-10; 5e-6 / 6e-7 after 64 steps on a
  bed 300 m above the datum;
- the arithmetic probe bit for bit against the CPU (two-sum error term, IEEE
  division and sqrt, the friction depth power, film drag), and the fast
  order's reciprocal against the IEEE division over its whole range.
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays
from inundation_amd.mesh.blocks import num_blocks
from oracle import cpu_ref, ref64
from tests import _drive as D
from tests._meshes import CASES, DRY_TOL, DT, MANNING, f32_state, step_shore, valley_grid
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():  # noqa: ANN201
    c = gpu_context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def grid():  # noqa: ANN201
    return valley_grid(64, 32, 2.0)


@pytest.fixture(scope="module")
def mesh(grid):  # noqa: ANN001, ANN201
    return device_mesh_arrays(grid)


def _both(ctx, mesh):  # noqa: ANN001, ANN202
    ctx.set_mesh(mesh["cell_l"], mesh["cell_r"], mesh["nx"], mesh["ny"], mesh["len"], mesh["dz2"],
                 mesh["area"], mesh["cptr"], mesh["cidx"])
    return ctx, cpu_ref.CpuRef(mesh)


def test_probe_is_bitwise_the_cpu(ctx) -> None:  # noqa: ANN001
    rng = np.random.default_rng(5)
    n = 1 << 18
    a = np.concatenate([
        rng.uniform(0.0, 2.0, n // 4),
        10.0 ** rng.uniform(-6, 3, n // 4),
        rng.normal(0, 100, n // 4),
        1.0 + rng.uniform(-1e-6, 1e-6, n // 4),
        # square roots across the whole range: subnormals, the library's
        # rescaled band below 2^-96, its edge, signed zeros
        (10.0 ** rng.uniform(-45.5, 38, n // 4)).astype(np.float32),
        np.float32(2.0 ** -96) * (1.0 + rng.uniform(-1e-3, 1e-3, 1024)).astype(np.float32),
        np.array([0.0, -0.0, 2.0 ** -149, 2.0 ** -126, 2.0 ** -96, 3.4e38], np.float32),
    ]).astype(np.float32)
    b = np.concatenate([
        10.0 ** rng.uniform(-9, -1, n // 2),
        rng.normal(0, 1, a.size - n // 2),
    ]).astype(np.float32)
    b[b == 0] = 1.0
    pairs = np.column_stack([a, b]).astype(np.float32)
    got = ctx.probe(pairs, 9.81, 1e-3)
    want = cpu_ref.probe(pairs, 9.81, 1e-3)
    for k, name in enumerate(["two_sum err", "div", "sqrt", "pow43", "film", "rcbrt", "fast film"]):
        both_nan = np.isnan(got[:, k]) & np.isnan(want[:, k])
        d = np.flatnonzero((got[:, k].view(np.uint32) != want[:, k].view(np.uint32)) & ~both_nan)
        assert d.size == 0, f"{name}: {d.size} differ, e.g. {pairs[d[:3]].tolist()} {got[d[:3], k]} {want[d[:3], k]}"
    # The two-sum error term survives: the reference's EFT probe (solver.py:453-461).
    e = ctx.probe(np.array([[1.0, 1e-9]], np.float32), 9.81, 1e-3)[0, 0]
    assert e == np.float32(1e-9)


def test_fast_reciprocal_is_the_ieee_division_on_its_range(ctx) -> None:  # noqa: ANN001
    """The fast order's rcp(x) (hardware estimate and one fused Newton step)
    is bitwise 1 / x for every float with 2^-126 <= |x| <= 2^126, both
    signs, about 4.2e9 cases; below that range (subnormals) and above it it
    is not, which is why every divisor it takes is kept inside."""
    n, first = ctx.probe_rcp(0x00800000, 0x7E800001)
    assert n == 0, f"{n} differ, first 0x{first:08x}"
    n_sub, first_sub = ctx.probe_rcp(0x00000001, 0x00800000)
    assert n_sub > 0 and first_sub == 0x00000001
    n_big, first_big = ctx.probe_rcp(0x7E800001, 0x7F800000)
    assert n_big > 0 and first_big == 0x7E800001


EXACT = L.FLAG_EXACT_ORDER
FLAG_SETS = {
    "comp+active": D.FLAGS_DEFAULT,
    "plain+active": L.FLAG_DETERMINISTIC | L.FLAG_ACTIVE_SET,
    "comp+all": L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED,
    "exact+comp+active": D.FLAGS_DEFAULT | EXACT,
    "exact+plain+all": L.FLAG_DETERMINISTIC | EXACT,
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("flags", list(FLAG_SETS.values()), ids=list(FLAG_SETS))
def test_fixed_dt_steps_are_bitwise_the_fp32_oracle(ctx, mesh, grid, case, flags) -> None:  # noqa: ANN001
    gpu, ref = _both(ctx, mesh)
    h, hu, hv = f32_state(*CASES[case](grid))
    nb = num_blocks(grid.N, 256)
    for be in (gpu, ref):
        D.setup(be, grid.N, h, hu, hv, MANNING, dry_tol=DRY_TOL, flags=flags)
        D.start(be, 8 * DT, DT, DT)
    D.assert_same(D.snapshot(gpu, grid.N, nb), D.snapshot(ref, grid.N, nb), "prologue")
    gpu.steps(3, use_graph=False)
    ref.steps(3)
    D.assert_same(D.snapshot(gpu, grid.N, nb), D.snapshot(ref, grid.N, nb), "3 steps")
    gpu.steps(7, use_graph=True)  # past the stop: the last two are no-ops
    ref.steps(7)
    got, want = D.snapshot(gpu, grid.N, nb), D.snapshot(ref, grid.N, nb)
    D.assert_same(got, want, "10 steps")
    assert int(got["ctrl"][1]) == 8 and int(got["ctrl"][0]) == 8 * 1024


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_fixed_dt_steps_match_fp64_within_reference_tolerance(ctx, mesh, grid, case, order) -> None:  # noqa: ANN001
    gpu, _ = _both(ctx, mesh)
    h, hu, hv = f32_state(*CASES[case](grid))
    D.setup(gpu, grid.N, h, hu, hv, MANNING, dry_tol=DRY_TOL, flags=D.FLAGS_DEFAULT | order)
    D.start(gpu, 8 * DT, DT, DT)
    gpu.steps(8, use_graph=True)
    got = D.snapshot(gpu, grid.N, num_blocks(grid.N, 256))
    m64 = ref64.Mesh64.from_device_arrays(mesh)
    s0 = ref64.State64(h.astype(np.float64), hu.astype(np.float64), hv.astype(np.float64))
    want = ref64.run_dt_sequence(m64, s0, np.full(grid.N, MANNING), [DT] * 8, 9.81, DRY_TOL)
    assert np.abs(got["h"] - want.h).max() <= 2e-6
    assert np.abs(got["hu"] - want.hu).max() <= 2e-5
    assert np.abs(got["hv"] - want.hv).max() <= 2e-5


@pytest.mark.parametrize("case", ["dam_break", "draining"])
@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_cfl_driven_run_is_bitwise_the_fp32_oracle(ctx, mesh, grid, case, order) -> None:  # noqa: ANN001
    """Every step picks its own dt on device; windows retarget the clock."""
    gpu, ref = _both(ctx, mesh)
    h, hu, hv = f32_state(*CASES[case](grid))
    nb = num_blocks(grid.N, 256)
    for be in (gpu, ref):
        D.setup(be, grid.N, h, hu, hv, MANNING, dry_tol=DRY_TOL, flags=D.FLAGS_DEFAULT | order)
        D.start(be, 2.0, 1.0, 0.01)
    for k in range(6):
        gpu.steps(25, use_graph=True)
        ref.steps(25)
        D.assert_same(D.snapshot(gpu, grid.N, nb), D.snapshot(ref, grid.N, nb), f"block {k}")
    # A second window: move the stop and re-plan.
    for be in (gpu, ref):
        c = be.ctrl_read()
        c.stop_ticks = int(c.stop_ticks) + 2 * 2**20
        be.ctrl_write(c)
        be.retarget()
    gpu.steps(50, use_graph=True)
    ref.steps(50)
    D.assert_same(D.snapshot(gpu, grid.N, nb), D.snapshot(ref, grid.N, nb), "after retarget")


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_sources_are_bitwise_the_fp32_oracle(ctx, mesh, grid, order) -> None:  # noqa: ANN001
    """A point source into a dry cell and rain on a patch of cells."""
    gpu, ref = _both(ctx, mesh)
    n = grid.N
    nb = num_blocks(n, 256)
    z = np.zeros(n, np.float32)
    x, y = grid.centroid[:, 0], grid.centroid[:, 1]
    point = int(np.argmin((x - 100.0) ** 2 + (y - 32.0) ** 2))
    patch = np.flatnonzero((x > 20.0) & (x < 60.0)).astype(np.uint32)
    cells = np.union1d(patch, [point]).astype(np.uint32)
    rates = np.where(cells == point, 0.5 / grid.area[point], 1.4e-5).astype(np.float32)
    for be in (gpu, ref):
        D.setup(be, n, z, z, z, MANNING, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates,
                flags=D.FLAGS_DEFAULT | order)
        D.start(be, 4.0, 0.5, 0.05)
    for k in range(4):
        gpu.steps(40, use_graph=True)
        ref.steps(40)
        D.assert_same(D.snapshot(gpu, n, nb), D.snapshot(ref, n, nb), f"block {k}")
    assert D.snapshot(gpu, n, nb)["h"].max() > DRY_TOL


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
@pytest.mark.parametrize("active", [True, False], ids=["active", "all"])
def test_sparse_rain_dt_bound_is_bitwise_the_fp32_oracle(ctx, mesh, grid, order, active) -> None:  # noqa: ANN001
    """Rain on every other cell beside a dam break, CFL-driven from a mostly dry start.

    Many wetted source cells have a dry neighbour in another block, the case
    whose dt bound is taken in two passes on the device; the clock and state
    must still follow the oracle bit for bit.
    """
    gpu, ref = _both(ctx, mesh)
    n = grid.N
    nb = num_blocks(n, 256)
    h, hu, hv = f32_state(*CASES["dam_break"](grid))
    cells = np.arange(0, n, 2, dtype=np.uint32)
    rates = np.where(cells % 6 == 0, 0.05, 0.004).astype(np.float32)
    flags = (D.FLAGS_DEFAULT if active else L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED) | order
    for be in (gpu, ref):
        D.setup(be, n, h, hu, hv, MANNING, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates, flags=flags)
        D.start(be, 3.0, 0.5, 0.05)
    for k in range(4):
        gpu.steps(30, use_graph=True)
        ref.steps(30)
        D.assert_same(D.snapshot(gpu, n, nb), D.snapshot(ref, n, nb), f"block {k}")


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_rain_on_a_wet_lake_is_bitwise_the_fp32_oracle(ctx, mesh, grid, order) -> None:  # noqa: ANN001
    """Rain on every cell of a moving lake: every block is wet after the
    sources, the case whose source dt bound the fast order takes from each
    cell's perimeter; the clock must still follow the oracle bit for bit."""
    gpu, ref = _both(ctx, mesh)
    n = grid.N
    nb = num_blocks(n, 256)
    h, hu, hv = f32_state(*CASES["wet_lake"](grid))
    cells = np.arange(n, dtype=np.uint32)
    rates = np.full(n, 2e-3, dtype=np.float32)
    rates[::7] = 0.0  # some cells without a source
    for be in (gpu, ref):
        D.setup(be, n, h, hu, hv, MANNING, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates,
                flags=D.FLAGS_DEFAULT | order)
        D.start(be, 3.0, 0.5, 0.05)
    for k in range(4):
        gpu.steps(30, use_graph=True)
        ref.steps(30)
        D.assert_same(D.snapshot(gpu, n, nb), D.snapshot(ref, n, nb), f"block {k}")


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_rain_beside_a_dry_bank_is_bitwise_the_fp32_oracle(ctx, order) -> None:  # noqa: ANN001
    """A block wet in every cell whose last row borders a dry block: its shore
    cells must take 2c across the dry edge, not the all-wet perimeter form."""
    g, h = step_shore()
    m = device_mesh_arrays(g)
    gpu, ref = _both(ctx, m)
    z = np.zeros(g.N, np.float32)
    cells = np.arange(g.N, dtype=np.uint32)
    rates = np.full(g.N, 2e-3, dtype=np.float32)
    nb = num_blocks(g.N, 256)
    for be in (gpu, ref):
        D.setup(be, g.N, h, z, z, MANNING, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates,
                flags=D.FLAGS_DEFAULT | order)
        D.start(be, 2.0, 0.5, 0.05)
    for k in range(3):
        gpu.steps(10, use_graph=True)
        ref.steps(10)
        D.assert_same(D.snapshot(gpu, g.N, nb), D.snapshot(ref, g.N, nb), f"block {k}")


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_sinks_that_dry_cells_are_bitwise_the_fp32_oracle(ctx, mesh, grid, order) -> None:  # noqa: ANN001
    """Negative source rates drain cells of a dam-break pool below dry_tol
    while rain falls elsewhere: the blocks' all-wet test and dry flags must
    follow the post-source depths, not the wet mark taken before the sinks."""
    gpu, ref = _both(ctx, mesh)
    n = grid.N
    nb = num_blocks(n, 256)
    h, hu, hv = f32_state(*CASES["dam_break"](grid))
    wet = np.flatnonzero(h > 0.0)
    sinks = wet[:: max(1, wet.size // 24)]
    cells = np.arange(n, dtype=np.uint32)
    rates = np.full(n, 2e-3, dtype=np.float32)
    rates[sinks] = -1.0
    for be in (gpu, ref):
        D.setup(be, n, h, hu, hv, MANNING, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates,
                flags=D.FLAGS_DEFAULT | order)
        D.start(be, 3.0, 0.05, 0.05)
    for k in range(3):
        gpu.steps(20, use_graph=True)
        ref.steps(20)
        D.assert_same(D.snapshot(gpu, n, nb), D.snapshot(ref, n, nb), f"block {k}")
    assert (gpu.download(L.FIELD_H, n)[sinks] < DRY_TOL).any()  # some sink cells ran dry


@pytest.fixture(scope="module")
def big():  # noqa: ANN201
    g = valley_grid(1024, 512, 2.0)
    return g, device_mesh_arrays(g)


def test_lists_and_source_fix_over_two_thousand_blocks(ctx, big) -> None:  # noqa: ANN001
    """2048 blocks: the active list spans several list-building workgroups,
    and rain on part of a dam break leaves blocks wet, dry and mixed, so the
    source fix skips some neighbour groups and works through others. Bitwise
    the oracle; the work counters equal the listed blocks' cells and slots."""
    g, m = big
    gpu, ref = _both(ctx, m)
    h, hu, hv = f32_state(*CASES["dam_break"](g))
    x = g.centroid[:, 0]
    cells = np.flatnonzero(x < 0.4 * x.max()).astype(np.uint32)
    rates = np.full(cells.size, 2e-3, dtype=np.float32)
    nb = num_blocks(g.N, 256)
    for be in (gpu, ref):
        D.setup(be, g.N, h, hu, hv, MANNING, dry_tol=DRY_TOL, src_cells=cells, src_rates=rates)
        D.start(be, 5.0, 0.5, 0.05)
    act = gpu.download(L.FIELD_BLOCK_ACTIVE, nb, np.uint32) != 0
    assert 0 < act.sum() < nb
    gpu.activity(reset=True)
    gpu.steps(1, use_graph=False)
    ref.steps(1)
    w = gpu.activity()
    cptr = m["cptr"].astype(np.int64)
    lo = np.arange(nb) * 256
    hi = np.minimum(lo + 256, g.N)
    assert w["steps"] == 1
    assert w["cells"] == int((hi - lo)[act].sum())
    assert w["slots"] == int((cptr[hi] - cptr[lo])[act].sum())
    for k in range(3):
        gpu.steps(20, use_graph=True)
        ref.steps(20)
        D.assert_same(D.snapshot(gpu, g.N, nb), D.snapshot(ref, g.N, nb), f"block {k}")
    # Active set off: every block listed, every cell, edge and slot counted.
    D.setup(gpu, g.N, h, hu, hv, MANNING, dry_tol=DRY_TOL, flags=L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED)
    D.start(gpu, 5.0, 0.5, 0.05)
    gpu.activity(reset=True)
    gpu.steps(4, use_graph=False)
    w = gpu.activity()
    assert (w["steps"], w["cells"], w["edges"], w["slots"]) == (4, 4 * g.N, 4 * g.E, 4 * int(cptr[g.N]))


def test_depths_uploaded_mid_run_reset_the_dry_flags(ctx, big) -> None:  # noqa: ANN001
    """A lake wet everywhere leaves every block's dry flag clear. Depths
    uploaded mid-run then empty the whole mesh, and rain falls on every cell
    of one block only. When its cells turn wet, every dry neighbour they have
    lies in another block that no step has looked at since the upload, so
    the dt bound (which sets the step here) is right only if the source fix
    takes those blocks as possibly dry, which the upload told it. Bitwise
    the oracle, clock included."""
    g, m = big
    gpu, ref = _both(ctx, m)
    n = g.N
    nb = num_blocks(n, 256)
    h, hu, hv = f32_state(*CASES["wet_lake"](g))
    for be in (gpu, ref):
        D.setup(be, n, h, hu, hv, MANNING, dry_tol=DRY_TOL)
        D.start(be, 1.0, 0.05, 0.05)
    gpu.steps(5, use_graph=False)
    ref.steps(5)
    z = np.zeros(n, dtype=np.float32)
    for which in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO):
        gpu.upload(which, z)
    ref.set_state(z, z, z, z)
    x, y = g.centroid[:, 0], g.centroid[:, 1]
    b = int(np.argmin((x - 0.8 * x.max()) ** 2 + (y - 0.5 * y.max()) ** 2)) // 256
    cells = np.arange(256 * b, min(256 * (b + 1), n), dtype=np.uint32)
    rates = np.full(cells.size, 5e-3, dtype=np.float32)
    for be in (gpu, ref):
        be.set_sources(cells, rates)
        D.start(be, 60.0, 5.0, 0.05)
    for k in range(4):
        gpu.steps(4, use_graph=False)
        ref.steps(4)
        D.assert_same(D.snapshot(gpu, n, nb), D.snapshot(ref, n, nb), f"after the upload, block {k}")
    c = gpu.ctrl_read()
    assert c.dt_ticks > 0.05 * 2**20  # the wetted block's bound took over from the dry-mesh dt


def test_active_set_on_equals_off(ctx, mesh, grid) -> None:  # noqa: ANN001
    """Skipping blocks is exact (reference test_gpu_active_set.py:77-91)."""
    gpu, _ = _both(ctx, mesh)
    h, hu, hv = f32_state(*CASES["dam_break"](grid))
    nb = num_blocks(grid.N, 256)
    out = []
    for flags in (D.FLAGS_DEFAULT, L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED):
        D.setup(gpu, grid.N, h, hu, hv, MANNING, dry_tol=DRY_TOL, flags=flags)
        D.start(gpu, 3.0, 1.0, 0.01)
        gpu.steps(150, use_graph=True)
        s = D.snapshot(gpu, grid.N, nb)
        out.append({k: s[k] for k in ("h", "hu", "hv", "h_lo", "ctrl")})
    D.assert_same(out[0], out[1], "active on vs off")


def test_stream_copy_reports_an_hbm_rate(ctx) -> None:  # noqa: ANN001
    ms = ctx.stream_copy(256 << 20, 3)
    gbs = 2 * (256 << 20) / (ms * 1e-3) / 1e9
    assert 500.0 < gbs < 8000.0
