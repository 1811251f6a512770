"""The benchmark workloads, built from seeds (seed = 1000 + workload number).

- ``lake_at_rest``: 64 x 64 quads of 10 m over 12 Gaussian bumps (some dry
  islands), still surface at 1 m, n = 0.03, CFL 0.5, 1000 steps;
- ``rainfall``: the 16 003 008-cell mixed triangle/quad lattice over 8 km of
  fBm terrain, 50 mm/h uniform rain from a dry start, per-cell n
  log-uniform in [0.015, 0.1], 2000 steps (``scale`` shrinks the lattice for
  tests);
- ``inflow``: a Lloyd-relaxed Voronoi mesh over 2 km of fBm terrain with a
  carved channel and one point source with a triangular hydrograph (peak
  20 m^3/s at 30 min over 1 h), n = 0.035;
- ``batched``: 64 scenarios over five such meshes (50 k to 2 M cells), each
  with its own inflow peak and rain;
- ``calibration``: ``inflow`` on 250 k cells with a smooth 'true' Manning
  field in [0.02, 0.08], 32 gauges, 2000 steps, for the adjoint.

Terrain is evaluated at the mean of each cell's vertices (the polygon centroid
for the lattice quads).
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

from .mesh.geometry import MeshGeometry, build_geometry, hilbert_reorder
from .mesh.synth import fbm_terrain, gaussian_bumps, lattice_tri_quad, quad_grid, sample_bilinear, voronoi_mesh
from .workflow import PointSource, SimulationPhase, build_phase_source_rate


@dataclass
class Workload:
    name: str
    geom: MeshGeometry
    h0: NDArray[np.float32]
    hu0: NDArray[np.float32]
    hv0: NDArray[np.float32]
    manning: NDArray[np.float32]
    source_rate: NDArray[np.float32] | None  # m^3/s per cell (constant)
    dry_tol: float
    cfl: float
    dt_max: float
    dt_init: float
    steps: int
    seed: int

    def source_pairs(self) -> tuple[NDArray[np.uint32], NDArray[np.float32]]:
        """(cell, dh/dt) pairs as the resident backend forms them (gpu_resident_batch.py:666-669)."""
        if self.source_rate is None:
            return np.empty(0, np.uint32), np.empty(0, np.float32)
        dh = np.asarray(self.source_rate, dtype=np.float32) * (1.0 / self.geom.area).astype(np.float32)
        cells = np.flatnonzero(dh).astype(np.uint32)
        return cells, dh[cells].astype(np.float32)


def _vertex_means(verts: NDArray[np.float64], faces: NDArray[np.int32], offsets: NDArray[np.int32]) -> tuple:
    deg = np.diff(offsets)
    cell = np.repeat(np.arange(deg.size), deg)
    cx = np.bincount(cell, verts[faces, 0], minlength=deg.size) / deg
    cy = np.bincount(cell, verts[faces, 1], minlength=deg.size) / deg
    return cx, cy


def lake_at_rest(seed: int = 1001) -> Workload:
    rng = np.random.default_rng(seed)
    nx = ny = 64
    dx = 10.0
    verts, faces = quad_grid(nx, ny, dx)
    cx = (np.arange(nx * ny) % nx + 0.5) * dx
    cy = (np.arange(nx * ny) // nx + 0.5) * dx
    zb = gaussian_bumps(cx, cy, nx * dx, rng)
    geom, _ = hilbert_reorder(build_geometry(verts, faces, zb_from_file=zb), verbose=False)
    h0 = np.maximum(np.float32(1.0) - geom.zb, np.float32(0.0)).astype(np.float32)
    z = np.zeros(geom.N, np.float32)
    return Workload(
        name="lake-at-rest 4096 quads (64x64 x 10 m, 12 Gaussian bumps), 1000 steps",
        geom=geom, h0=h0, hu0=z, hv0=z.copy(), manning=np.full(geom.N, 0.03, np.float32),
        source_rate=None, dry_tol=1e-4, cfl=0.5, dt_max=1.0, dt_init=0.01, steps=1000, seed=seed,
    )


def rainfall(scale: int = 1, seed: int = 1003, rain_mm_h: float = 50.0) -> Workload:
    """The 16 M-cell rainfall flood; ``scale`` k uses a 3024/k lattice over 8000/k m."""
    rng = np.random.default_rng(seed)
    nq = 3024 // scale
    length = 8000.0 / scale
    verts, faces, offsets = lattice_tri_quad(nq, length, rng)
    cx, cy = _vertex_means(verts, faces, offsets)
    zb = fbm_terrain(cx, cy, length, rng)
    del cx, cy
    n_cells = offsets.size - 1
    manning = np.exp(rng.uniform(np.log(0.015), np.log(0.1), n_cells)).astype(np.float32)
    geom = build_geometry(verts, faces, offsets, zb_from_file=zb, manning_n_from_file=manning)
    del verts, faces, offsets, zb
    geom, _ = hilbert_reorder(geom, verbose=False)
    # Uniform rain as one domain-covering point source: the area-proportional
    # split gives every cell the same intensity (reference workflow.py:617-624).
    total_area = float(np.float64(geom.area.astype(np.float64).sum()))
    q = rain_mm_h / 1000.0 / 3600.0 * total_area
    phase = SimulationPhase(1800.0, [PointSource(q, (length / 2, length / 2), 2.0 * length)])
    src = build_phase_source_rate(geom, phase, geom.area > 1e-6)
    z = np.zeros(geom.N, np.float32)
    return Workload(
        name=f"rainfall {geom.N} mixed tri/quad cells over {length:g} m fBm, {rain_mm_h:g} mm/h from dry",
        geom=geom, h0=z, hu0=z.copy(), hv0=z.copy(), manning=geom.manning_n, source_rate=src,
        dry_tol=1e-4, cfl=0.8, dt_max=0.5, dt_init=0.5, steps=2000, seed=seed,
    )


def inflow(
    n_points: int = 1_000_000, seed: int = 1002, length: float = 2000.0, processes: int = 8
) -> tuple[Workload, list[SimulationPhase]]:
    """Voronoi inflow flood and its 60 one-minute hydrograph phases (``processes``: mesh-building workers)."""
    rng = np.random.default_rng(seed)
    verts, faces, offsets = voronoi_mesh(n_points, length, rng, processes=processes)
    cx, cy = _vertex_means(verts, faces, offsets)
    zb = fbm_terrain(cx, cy, length, rng)
    geom, _ = hilbert_reorder(build_geometry(verts, faces, offsets, zb_from_file=zb), verbose=False)
    # The source sits in the channel near the upstream edge.
    x0 = 0.05 * length
    y0 = 0.5 * length + 0.1 * length * np.sin(2.0 * np.pi * x0 / length)
    phases = []
    for k in range(60):
        t_mid = (k + 0.5) * 60.0
        q = 20.0 * (t_mid / 1800.0 if t_mid <= 1800.0 else (3600.0 - t_mid) / 1800.0)
        phases.append(SimulationPhase(60.0, [PointSource(q, (x0, y0), 0.0)]))
    z = np.zeros(geom.N, np.float32)
    wl = Workload(
        name=f"inflow {geom.N} Voronoi cells over {length:g} m, 20 m^3/s peak hydrograph",
        geom=geom, h0=z, hu0=z.copy(), hv0=z.copy(), manning=np.full(geom.N, 0.035, np.float32),
        source_rate=None, dry_tol=1e-4, cfl=0.8, dt_max=1.0, dt_init=0.1, steps=0, seed=seed,
    )
    return wl, phases


def _voronoi_terrain(n_points: int, seed: int, length: float, processes: int = 1) -> MeshGeometry:
    rng = np.random.default_rng(seed)
    verts, faces, offsets = voronoi_mesh(n_points, length, rng, processes=processes)
    cx, cy = _vertex_means(verts, faces, offsets)
    zb = fbm_terrain(cx, cy, length, rng)
    return hilbert_reorder(build_geometry(verts, faces, offsets, zb_from_file=zb), verbose=False)[0]


@dataclass
class Scenario:
    """One flood of the batched workload: its mesh, inflow peak and rain."""

    mesh: int
    peak: float  # m^3/s at 30 min, triangular over 1 h
    rain_mm_h: float
    inflow_cell: int

    def inflow(self, t_mid: float) -> float:
        """Discharge of the one-minute phase centred on t_mid (s)."""
        return self.peak * (t_mid / 1800.0 if t_mid <= 1800.0 else max(3600.0 - t_mid, 0.0) / 1800.0)

    def source_pairs(self, geom: MeshGeometry, t_mid: float) -> tuple[NDArray[np.uint32], NDArray[np.float32]]:
        """(cell, dh/dt) of the phase: uniform rain on every cell plus the inflow cell."""
        rain = np.float32(self.rain_mm_h / 1000.0 / 3600.0)
        if rain > 0.0:
            cells = np.arange(geom.N, dtype=np.uint32)
            rates = np.full(geom.N, rain, np.float32)
        else:
            cells = np.array([self.inflow_cell], np.uint32)
            rates = np.zeros(1, np.float32)
        k = int(np.searchsorted(cells, self.inflow_cell))
        rates[k] = np.float32(rates[k] + np.float32(self.inflow(t_mid)) / geom.area[self.inflow_cell])
        return cells, rates


def batched(
    n_scenarios: int = 64,
    sizes: tuple[int, ...] = (50_000, 200_000, 500_000, 1_000_000, 2_000_000),
    seed: int = 1004,
    length: float = 2000.0,
    processes: int = 5,
    progress: Callable[[str], None] | None = None,
) -> tuple[list[MeshGeometry], list[Scenario]]:
    """The batched-scenario workload: Voronoi meshes (terrain as ``inflow``) and the scenarios on them.

    Scenario k runs on mesh k mod len(sizes), with a seeded inflow peak
    U[5, 50] m^3/s into the channel near the upstream edge and uniform rain
    U[0, 80] mm/h. With ``processes`` > 1 the meshes are built in spawned
    worker processes, safe in a process that already uses a GPU.
    """
    import multiprocessing as mp

    args = [(n, seed * 10 + i, length) for i, n in enumerate(sizes)]
    geoms: list[MeshGeometry | None] = [None] * len(sizes)
    if processes > 1 and len(sizes) > 1:
        # The largest mesh is built here with its Lloyd tiles spread over
        # workers, the others side by side in a pool (spawned workers: safe
        # in a process that already uses a GPU).
        big = int(np.argmax(sizes))
        rest = [k for k in range(len(sizes)) if k != big]
        with mp.get_context("spawn").Pool(min(processes, len(rest))) as pool:
            pending = {k: pool.apply_async(_voronoi_terrain, args[k]) for k in rest}
            geoms[big] = _voronoi_terrain(*args[big], processes=max(2, processes))
            if progress:
                progress(f"mesh of {sizes[big]} points built")
            for k in rest:
                geoms[k] = pending[k].get()
                if progress:
                    progress(f"mesh of {sizes[k]} points built")
    else:
        for k, x in enumerate(args):
            geoms[k] = _voronoi_terrain(*x)
            if progress:
                progress(f"mesh of {x[0]} points built")
    x0 = 0.05 * length
    y0 = 0.5 * length + 0.1 * length * np.sin(2.0 * np.pi * x0 / length)
    inlet = [int(np.argmin((g.centroid[:, 0] - x0) ** 2 + (g.centroid[:, 1] - y0) ** 2)) for g in geoms]
    scen = []
    for k in range(n_scenarios):
        rng = np.random.default_rng([seed, k])
        m = k % len(sizes)
        scen.append(Scenario(mesh=m, peak=float(rng.uniform(5.0, 50.0)), rain_mm_h=float(rng.uniform(0.0, 80.0)),
                             inflow_cell=inlet[m]))
    return geoms, scen


def smooth_field(
    x: NDArray[np.float64],
    y: NDArray[np.float64],
    length: float,
    rng: np.random.Generator,
    lo: float,
    hi: float,
    sigma: float = 150.0,
    grid: int = 256,
) -> NDArray[np.float32]:
    """Gaussian-filtered white noise (``sigma`` metres) mapped linearly onto [lo, hi] over the cells."""
    noise = rng.standard_normal((grid, grid))
    k = np.fft.fftfreq(grid, d=length / grid)
    gain = np.exp(-2.0 * np.pi**2 * sigma**2 * (k[:, None] ** 2 + k[None, :] ** 2))
    f = np.fft.ifft2(np.fft.fft2(noise) * gain).real
    v = sample_bilinear(f, length, x, y)
    v = (v - v.min()) / max(float(v.max() - v.min()), 1e-300)
    return (lo + (hi - lo) * v).astype(np.float32)


@dataclass
class Calibration:
    """Manning calibration: a 'true' n field, gauges, and the inflow of ``inflow`` on a Voronoi mesh."""

    geom: MeshGeometry
    n_true: NDArray[np.float32]
    inflow_cell: int
    steps: int = 2000
    seg: int = 100
    every: int = 10
    n_gauges: int = 32
    n_start: float = 0.04
    seed: int = 1005
    peak: float = 20.0

    def inflow(self, t: float) -> float:
        """Discharge (m^3/s) of the one-minute hydrograph phase holding time t (s)."""
        t_mid = (np.floor(t / 60.0) + 0.5) * 60.0
        return self.peak * (t_mid / 1800.0 if t_mid <= 1800.0 else max(3600.0 - t_mid, 0.0) / 1800.0)

    def source_cells(self) -> NDArray[np.uint32]:
        return np.array([self.inflow_cell], np.uint32)

    def rates_at(self, t: float) -> NDArray[np.float32]:
        """dh/dt of the source cell in the phase holding t."""
        return np.array([np.float32(self.inflow(t)) / self.geom.area[self.inflow_cell]], np.float32)

    def pick_gauges(self, h_truth: NDArray[np.float32], min_depth: float = 0.05) -> NDArray[np.uint32]:
        """``n_gauges`` seeded cells among those deeper than ``min_depth`` at the end of the truth run."""
        wet = np.flatnonzero(h_truth > min_depth)
        if wet.size < self.n_gauges:
            raise ValueError(f"only {wet.size} cells are wet in the truth run")
        rng = np.random.default_rng([self.seed, 1])
        return np.sort(rng.choice(wet, self.n_gauges, replace=False)).astype(np.uint32)


def calibration(n_points: int = 250_000, seed: int = 1005, length: float = 2000.0, steps: int = 2000) -> Calibration:
    """The adjoint calibration workload: ``inflow``'s terrain and hydrograph on ``n_points`` Voronoi cells.

    The true Manning field is seeded smooth noise (150 m Gaussian filter)
    mapped onto [0.02, 0.08]; the calibration starts from n = 0.04.
    """
    geom = _voronoi_terrain(n_points, seed, length)
    rng = np.random.default_rng([seed, 0])
    cx = geom.centroid[:, 0].astype(np.float64)
    cy = geom.centroid[:, 1].astype(np.float64)
    n_true = smooth_field(cx, cy, length, rng, 0.02, 0.08)
    x0 = 0.05 * length
    y0 = 0.5 * length + 0.1 * length * np.sin(2.0 * np.pi * x0 / length)
    inlet = int(np.argmin((cx - x0) ** 2 + (cy - y0) ** 2))
    return Calibration(geom=geom, n_true=n_true, inflow_cell=inlet, steps=steps, seed=seed)


__all__ = ["Calibration", "Scenario", "Workload", "batched", "calibration", "inflow", "lake_at_rest", "rainfall",
           "smooth_field"]
