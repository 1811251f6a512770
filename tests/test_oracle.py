"""The oracles pinned against the reference's own FP64 code and closed forms.

- ``oracle/ref64.py`` against fixtures written by the reference's FP64 test
  oracle ``tests/unit/_cpu_step.py`` (8-step trajectories of the four
  step-parity cases on a valley 300 m above the datum, the CFL bound, HLLC
  fluxes of random face states), to float64 rounding;
- ``oracle/cpu_ref.cpp`` (FP32, both arithmetic orders) against the same
  trajectories within the reference's FP32 tolerances (2e-6 m depth, 2e-5
  momentum; ``tests/unit/test_gpu_step_parity.py:31-35``);
- HLLC wave-speed known answers (``tests/unit/test_hllc_speeds.py:27-51``);
- planted mistakes in the FP64 oracle (a dropped bed step, a lost dry-bed
  speed, swapped normal components) fail the fixtures, so they have teeth.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from inundation_amd.hip.solver import device_mesh_arrays
from inundation_amd.mesh.geometry import build_geometry, hilbert_reorder
from oracle import cpu_ref, ref64
from tests import _drive as D

GOLD = Path(__file__).resolve().parent / "golden"
CASES = ("wet_lake", "dam_break", "draining", "dry_momentum")
DT = 2.0**-10
G = 9.81


@pytest.fixture(scope="module")
def valley():  # noqa: ANN201
    d = np.load(GOLD / "geometry_valley32x16.npz")
    g, _ = hilbert_reorder(build_geometry(d["in_verts"], d["in_faces"], zb_from_file=d["in_zb"]), verbose=False)
    return g


@pytest.fixture(scope="module")
def steps():  # noqa: ANN201
    return np.load(GOLD / "cpu_step_valley.npz")


def _run64(geom, steps, case):  # noqa: ANN001, ANN202
    m = ref64.Mesh64.from_device_arrays(device_mesh_arrays(geom))
    s0 = ref64.State64(*(steps[f"{case}_{k}0"].astype(np.float64) for k in ("h", "hu", "hv")))
    return ref64.run_dt_sequence(m, s0, np.full(geom.N, 0.03), [DT] * 8, G, 1e-3)


@pytest.mark.parametrize("case", CASES)
def test_fp64_oracle_matches_the_reference_fp64_step(valley, steps, case) -> None:  # noqa: ANN001
    s = _run64(valley, steps, case)
    for k, got in (("h", s.h), ("hu", s.hu), ("hv", s.hv)):
        np.testing.assert_allclose(got, steps[f"{case}_{k}8"], rtol=0, atol=1e-10, err_msg=k)


@pytest.mark.parametrize("case", CASES)
def test_fp64_cfl_bound_matches_the_reference(valley, steps, case) -> None:  # noqa: ANN001
    m = ref64.Mesh64.from_device_arrays(device_mesh_arrays(valley))
    s0 = ref64.State64(*(steps[f"{case}_{k}0"].astype(np.float64) for k in ("h", "hu", "hv")))
    assert ref64.prologue_bound(m, s0, G, 1e-3, 0.8) == pytest.approx(float(steps[f"{case}_cfl"]), rel=1e-12)


def test_fp64_hllc_matches_the_reference_vectors() -> None:
    d = np.load(GOLD / "hllc_vectors.npz")
    Fh, Fhn, Fht, smax = ref64.hllc(d["hL"], d["qnL"], d["qtL"], d["hR"], d["qnR"], d["qtR"], G, float(d["dry"]))
    for got, k in ((Fh, "Fh"), (Fhn, "Fhn"), (Fht, "Fht"), (smax, "smax")):
        np.testing.assert_allclose(got, d[k], rtol=1e-13, atol=1e-15, err_msg=k)
    SL, SR = ref64.hllc_speeds(d["hL"], d["unL"], d["hR"], d["unR"], G, float(d["dry"]))
    np.testing.assert_allclose(SL, d["SL"], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(SR, d["SR"], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("order", [0, 16], ids=["fast", "exact"])
def test_fp32_oracle_is_within_the_reference_tolerance(valley, steps, case, order) -> None:  # noqa: ANN001
    ref = cpu_ref.CpuRef(device_mesh_arrays(valley))
    h, hu, hv = (steps[f"{case}_{k}0"] for k in ("h", "hu", "hv"))
    D.setup(ref, valley.N, h, hu, hv, 0.03, dry_tol=1e-3, flags=D.FLAGS_DEFAULT | order)
    D.start(ref, 8 * DT, DT, DT)
    ref.steps(10)
    out = D.snapshot(ref, valley.N, 2)
    assert int(out["ctrl"][1]) == 8
    assert np.abs(out["h"] - steps[f"{case}_h8"]).max() <= 2e-6
    assert np.abs(out["hu"] - steps[f"{case}_hu8"]).max() <= 2e-5
    assert np.abs(out["hv"] - steps[f"{case}_hv8"]).max() <= 2e-5


# ---- HLLC speed known answers (closed forms) ---------------------------------

TOL = 1e-3


def _speeds(h_l: float, u_l: float, h_r: float, u_r: float) -> tuple[float, float]:
    sl, sr = ref64.hllc_speeds(np.array([h_l]), np.array([u_l]), np.array([h_r]), np.array([u_r]), G, TOL)
    return float(sl[0]), float(sr[0])


@pytest.mark.parametrize("edge", [TOL, 2.0 * TOL])
@pytest.mark.parametrize("wet", [(0.05, 0.3), (0.5, -0.4), (0.01, 0.0)])
def test_speeds_are_continuous_as_a_side_dries(edge: float, wet: tuple[float, float]) -> None:
    h, u = wet
    eps = 1e-9
    assert _speeds(h, u, edge * (1 - eps), 0.0) == pytest.approx(_speeds(h, u, edge * (1 + eps), 0.0), abs=1e-5)
    assert _speeds(edge * (1 - eps), 0.0, h, u) == pytest.approx(_speeds(edge * (1 + eps), 0.0, h, u), abs=1e-5)


def test_a_dry_side_gets_the_dry_bed_speed() -> None:
    h, u = 0.5, 0.3
    c = float(np.sqrt(G * h))
    assert _speeds(h, u, 0.0, 0.0) == pytest.approx((u - c, u + 2.0 * c))
    assert _speeds(0.0, 0.0, h, -u) == pytest.approx((-u - 2.0 * c, -u + c))


def test_two_wet_sides_keep_the_roe_estimate() -> None:
    hl, hr = 0.5, 0.2
    sl, sr = _speeds(hl, 0.0, hr, 0.0)
    c_roe = float(np.sqrt(G * 0.5 * (hl + hr)))
    assert sl == pytest.approx(min(-float(np.sqrt(G * hl)), -c_roe))
    assert sr == pytest.approx(max(float(np.sqrt(G * hr)), c_roe))


def test_still_water_flux_is_the_hydrostatic_pressure() -> None:
    """Equal still depths: no mass flux, normal momentum flux g h^2 / 2."""
    h = np.array([0.3, 1.7, 4.0])
    z = np.zeros_like(h)
    Fh, Fhn, Fht, _ = ref64.hllc(h, z, z, h, z, z, G, TOL)
    np.testing.assert_allclose(Fh, 0.0, atol=1e-15)
    np.testing.assert_allclose(Fhn, 0.5 * G * h * h, rtol=1e-14)
    np.testing.assert_allclose(Fht, 0.0, atol=1e-15)


def test_flux_is_antisymmetric_under_mirroring() -> None:
    """Swapping the sides and reversing the normal negates the mass flux."""
    rng = np.random.default_rng(3)
    hl, hr = rng.uniform(0.01, 2, 50), rng.uniform(0.01, 2, 50)
    ql, qr = rng.normal(0, 0.5, 50) * hl, rng.normal(0, 0.5, 50) * hr
    z = np.zeros(50)
    f1 = ref64.hllc(hl, ql, z, hr, qr, z, G, TOL)
    f2 = ref64.hllc(hr, -qr, z, hl, -ql, z, G, TOL)
    np.testing.assert_allclose(f1[0], -f2[0], atol=1e-12)
    np.testing.assert_allclose(f1[1], f2[1], atol=1e-12)


def test_lake_at_rest_over_bumps_stays_at_rest_in_fp64() -> None:
    """Audusse reconstruction is well balanced: still water over uneven bed barely moves.

    The surface is exactly 1 m in float64 (h = 1 - z from the float32 bed). The
    bed steps are stored as float32 (as in the reference, geometry.py:241-254),
    so where two neighbouring beds are far apart in magnitude the step is
    rounded by up to 3e-8 m; that is the only residual, and it moves the water
    by less than 1e-8 m/s.
    """
    from inundation_amd.workloads import lake_at_rest

    wl = lake_at_rest()
    m = ref64.Mesh64.from_device_arrays(device_mesh_arrays(wl.geom))
    h0 = np.maximum(1.0 - wl.geom.zb.astype(np.float64), 0.0)
    assert (h0 == 0.0).any(), "the bumps should leave dry islands"
    s = ref64.State64(h0, np.zeros(wl.geom.N), np.zeros(wl.geom.N))
    s = ref64.run_dt_sequence(m, s, wl.manning.astype(np.float64), [0.5] * 20, G, wl.dry_tol)
    wet = s.h > 10 * wl.dry_tol
    assert np.abs(s.hu[wet] / s.h[wet]).max() < 1e-8
    assert np.abs(s.hv[wet] / s.h[wet]).max() < 1e-8


# ---- planted mistakes -------------------------------------------------------


def _fails(valley, steps, monkeypatch, name, fn) -> bool:  # noqa: ANN001
    monkeypatch.setattr(ref64, name, fn)
    try:
        for case in CASES:
            s = _run64(valley, steps, case)
            for k, got in (("h", s.h), ("hu", s.hu), ("hv", s.hv)):
                if np.abs(got - steps[f"{case}_{k}8"]).max() > 1e-10:
                    return True
        return False
    finally:
        monkeypatch.undo()


def test_planted_mistakes_fail_the_fixtures(valley, steps, monkeypatch) -> None:  # noqa: ANN001
    face = ref64.face_states
    speeds = ref64.hllc_speeds
    edge_words = ref64.edge_words

    def no_bed_step(m, hL, huL, hvL, hR, huR, hvR, dry):  # noqa: ANN001, ANN202
        m2 = ref64.Mesh64(m.cl, m.cr, m.nx, m.ny, m.length, 0 * m.dzl, 0 * m.dzr, m.area, m.cptr, m.cidx)
        return face(m2, hL, huL, hvL, hR, huR, hvR, dry)

    def roe_only(hL, unL, hR, unR, g, dry):  # noqa: ANN001, ANN202
        return speeds(hL, unL, hR, unR, g, 1e-300)

    def swapped_normal(m, s, g, dry):  # noqa: ANN001, ANN202
        m2 = ref64.Mesh64(m.cl, m.cr, m.ny, m.nx, m.length, m.dzl, m.dzr, m.area, m.cptr, m.cidx)
        return edge_words(m2, s, g, dry)

    assert _fails(valley, steps, monkeypatch, "face_states", no_bed_step)
    assert _fails(valley, steps, monkeypatch, "hllc_speeds", roe_only)
    assert _fails(valley, steps, monkeypatch, "edge_words", swapped_normal)
    assert not _fails(valley, steps, monkeypatch, "edge_words", edge_words)
