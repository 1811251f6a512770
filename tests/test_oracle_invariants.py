"""Invariants of the FP32 oracle's resident run (the bitwise comparator of the GPU).

Each restates a property the reference tests on its GPU backend:

- lake at rest on the 4096-quad bump field (workload 1): max |u| over cells
  deeper than 10 dry_tol after 1000 steps at CFL 0.5 stays below 1e-6 m/s in
  the default (fast) arithmetic; the metric is the reference's
  (``inundation/bench/lake_at_rest.py:197-212``);
- an exactly level surface stays bitwise still in the fast arithmetic;
- a closed basin conserves volume: compensated depth within 1e-9 relative
  (``tests/unit/test_gpu_compensated.py:95-116``, which asks < 3e-9);
- a point source injects Q t (``tests/unit/test_gpu_resident_clock.py:284-314``);
- skipping inactive blocks is exact and repeat runs are bit-identical
  (``tests/unit/test_gpu_active_set.py:77-91``, ``test_gpu_deterministic.py:94-100``);
- dry cells carry no momentum (``tests/unit/test_gpu_step_parity.py:187-192``);
- the compensated pair keeps |lo| <= ulp(h)/2 (``test_gpu_compensated.py:266-305``).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from inundation_amd.hip.solver import device_mesh_arrays
from inundation_amd.mesh.geometry import build_geometry, edge_bed_steps
from inundation_amd.mesh.synth import quad_grid
from inundation_amd.workloads import lake_at_rest
from oracle import cpu_ref, ref64
from tests import _drive as D
from tests._meshes import CASES, f32_state, step_shore, valley_grid

EXACT = 16


def _max_speed(s: dict, dry_tol: float) -> float:
    h, hu, hv = s["h"], s["hu"], s["hv"]
    wet = h > 10 * dry_tol
    return float(max(np.abs(hu[wet] / h[wet]).max(), np.abs(hv[wet] / h[wet]).max()))


def _run(geom, h, hu, hv, n, steps, *, flags=D.FLAGS_DEFAULT, dry_tol=1e-4, cfl=0.8, dt_max=1.0, dt_init=0.01,
         src=None, stop=1e5):  # noqa: ANN001, ANN202
    ref = cpu_ref.CpuRef(device_mesh_arrays(geom))
    cells, rates = (None, None) if src is None else src
    D.setup(ref, geom.N, h, hu, hv, n, dry_tol=dry_tol, cfl=cfl, flags=flags, src_cells=cells, src_rates=rates)
    D.start(ref, stop, dt_max, dt_init)
    ref.steps(steps)
    return ref, D.snapshot(ref, geom.N, -(-geom.N // 256))


def test_lake_at_rest_stays_below_1e6_m_per_s() -> None:
    wl = lake_at_rest()
    _, s = _run(wl.geom, wl.h0, wl.hu0, wl.hv0, wl.manning, 1000, dry_tol=wl.dry_tol, cfl=wl.cfl,
                dt_max=wl.dt_max, dt_init=wl.dt_init)
    assert int(s["ctrl"][1]) == 1000
    assert _max_speed(s, wl.dry_tol) < 1e-6


def test_level_surface_is_bitwise_still() -> None:
    """Beds on a 2^-23 grid make 1 - z exact: nothing may move, not even the low words."""
    wl = lake_at_rest()
    g = wl.geom
    zs = (np.round(g.zb.astype(np.float64) * 2**23) / 2**23).astype(np.float32)
    dzl, dzr = edge_bed_steps(zs, g.edge_cellL, g.edge_cellR)
    g2 = dataclasses.replace(g, zb=zs, edge_dzL=dzl, edge_dzR=dzr)
    h0 = np.maximum(np.float32(1.0) - zs, np.float32(0.0)).astype(np.float32)
    _, s = _run(g2, h0, wl.hu0, wl.hv0, wl.manning, 300, dry_tol=wl.dry_tol, cfl=wl.cfl, dt_max=1.0)
    assert np.array_equal(s["h"], h0)
    assert not s["hu"].any() and not s["hv"].any() and not s["h_lo"].any()


def _basin(n: int = 24):  # noqa: ANN202
    verts, faces = quad_grid(n, n, 1.0)
    return build_geometry(verts, faces, zb_from_file=np.zeros(n * n, dtype=np.float32))


def _hump(g) -> np.ndarray:  # noqa: ANN001
    cx, cy = g.centroid[:, 0], g.centroid[:, 1]
    return (0.5 + 0.2 * np.exp(-((cx - 12.0) ** 2 + (cy - 12.0) ** 2) / 10.0)).astype(np.float32)


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_closed_basin_conserves_volume(order: int) -> None:
    g = _basin()
    h0 = _hump(g)
    z = np.zeros(g.N, np.float32)
    _, s = _run(g, h0, z, z, 0.03, 400, flags=D.FLAGS_DEFAULT | order, dt_max=1.0)
    v0 = float((h0.astype(np.float64) * g.area).sum())
    v1 = float(((s["h"].astype(np.float64) + s["h_lo"]) * g.area).sum())
    assert abs(v1 - v0) / v0 < 1e-9
    lo, h = s["h_lo"].astype(np.float64), s["h"]
    assert (np.abs(lo) <= 0.5 * np.spacing(h).astype(np.float64)).all()


def test_point_source_injects_q_times_t() -> None:
    g = _basin()
    h0 = np.full(g.N, 0.05, np.float32)
    z = np.zeros(g.N, np.float32)
    cell = g.N // 2 + 12
    q = 1.0
    src = (np.array([cell], np.uint32), np.array([q / g.area[cell]], np.float32))
    _, s = _run(g, h0, z, z, 0.03, 100000, src=src, stop=30.0, dt_max=1.0)
    v0 = float((h0.astype(np.float64) * g.area).sum())
    v1 = float(((s["h"].astype(np.float64) + s["h_lo"]) * g.area).sum())
    assert int(s["ctrl"][0]) == 30 * 2**20
    assert v1 - v0 == pytest.approx(q * 30.0, rel=1e-4)


@pytest.mark.parametrize("case", ["dam_break", "draining"])
def test_active_set_on_equals_off_and_runs_repeat(case: str) -> None:
    g = valley_grid(256, 128, 2.0)
    h, hu, hv = f32_state(*CASES[case](g))
    kw = dict(dry_tol=1e-3, dt_max=1.0)
    _, on = _run(g, h, hu, hv, 0.03, 120, **kw)
    _, again = _run(g, h, hu, hv, 0.03, 120, **kw)
    _, off = _run(g, h, hu, hv, 0.03, 120, flags=D.FLAGS_DEFAULT & ~4, **kw)
    for k in ("h", "hu", "hv", "h_lo", "ctrl"):
        assert np.array_equal(on[k], again[k]), k
        assert np.array_equal(on[k], off[k]), k
    assert on["active"].sum() < on["active"].size  # some blocks were skipped
    dry = on["h"] < 1e-3
    assert not on["hu"][dry].any() and not on["hv"][dry].any()


def test_dry_cells_start_and_stay_without_momentum() -> None:
    g = valley_grid(64, 32, 2.0)
    h, hu, hv = f32_state(*CASES["dry_momentum"](g))
    _, s = _run(g, h, hu, hv, 0.03, 8, dry_tol=1e-3, dt_max=2.0**-10, dt_init=2.0**-10, stop=8 * 2.0**-10)
    dry = s["h"] < 1e-3
    assert not s["hu"][dry].any() and not s["hv"][dry].any()


def _first_step_bounds(g, h, hu, hv, rates: np.ndarray, order: int) -> tuple[float, float, float]:  # noqa: ANN001
    """The FP32 oracle's next-step CFL bound after one step with rain, and the
    FP64 oracle's bound after the same step with and without the sources."""
    n = g.N
    cells = np.flatnonzero(rates).astype(np.uint32)
    ref, _ = _run(g, h, hu, hv, 0.03, 0, flags=D.FLAGS_DEFAULT | order, src=(cells, rates[cells]),
                  dry_tol=1e-3, dt_max=0.5, dt_init=0.05)
    dt = int(ref.ctrl_read().dt_ticks) / 2**20
    ref.steps(1)
    got = float(ref.ctrl_read().last_raw)
    m64 = ref64.Mesh64.from_device_arrays(device_mesh_arrays(g))
    s0 = ref64.State64(*(a.astype(np.float64) for a in (h, hu, hv)))
    man = np.full(n, 0.03)
    _, with_src = ref64.heun_step(m64, s0, man, dt, 9.81, 1e-3, 0.8, rates.astype(np.float64))
    _, no_src = ref64.heun_step(m64, s0, man, dt, 9.81, 1e-3, 0.8)
    return got, with_src, no_src


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_rain_on_a_wet_lake_source_bound_is_the_fp64_one(order: int) -> None:
    """Rain on every cell of a moving lake: each block is wet after the
    sources, where the fast order sums the bound over the cell perimeter
    (resident.py:709-728 sums (|u| + c) len edge by edge). The source bound
    is the binding one here, and dropping its |u| term moves it by more than
    the tolerance, so a wrong perimeter or a lost term fails."""
    g = valley_grid(64, 32, 2.0)
    h, hu, hv = f32_state(*CASES["wet_lake"](g))
    rates = np.full(g.N, 2e-3, dtype=np.float32)
    rates[::7] = 0.0
    got, with_src, no_src = _first_step_bounds(g, h, hu, hv, rates, order)
    assert with_src < no_src * (1.0 - 1e-4)  # the sources set the bound
    assert got == pytest.approx(with_src, rel=1e-5)
    # |u| is up to 0.1 m/s against c of about 2.5 m/s: a bound without it is
    # several 1e-3 larger.
    i = np.flatnonzero(rates)
    perim = np.add.reduceat(g.edge_len[g.cell_edge_idx].astype(np.float64), g.cell_edge_ptr[:-1])
    c_only = float(np.min(0.8 * g.area[i] / (np.sqrt(9.81 * h[i].astype(np.float64)) * perim[i])))
    assert abs(c_only - got) / got > 1e-3


@pytest.mark.parametrize("order", [0, EXACT], ids=["fast", "exact"])
def test_rain_beside_a_dry_bank_source_bound_is_the_fp64_one(order: int) -> None:
    """A block wet in every cell whose cells border a dry block: the shore
    cells take 2c across the dry edge (resident.py:709-728), which sets the
    bound at 4/5 of the edge-free one. The perimeter form must not apply."""
    g, h = step_shore()
    z = np.zeros(g.N, np.float32)
    rates = np.full(g.N, 2e-3, dtype=np.float32)
    got, with_src, no_src = _first_step_bounds(g, h, z, z, rates, order)
    assert with_src == pytest.approx(0.8 * no_src, rel=1e-3)
    assert got == pytest.approx(with_src, rel=1e-5)
