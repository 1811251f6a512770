"""Application API: ``SWEWorkflow`` with the ``hip_resident`` backend.

Mirrors the reference's ``inundation/workflow.py``: the frozen
``WorkflowConfig`` (``:135-227``), ``PointSource`` (``:60-71``),
``SimulationPhase`` (``:123-132``), ``WorkflowResult`` (``:230-255``), the
area-proportional point-source split (``:317-358, 599-625``), initial-state
loading (``:450-596``) and ``SWEWorkflow.prepare()/run()`` (``:658-969``).

``solver_impl`` accepts the reference's names and runs each on the HIP
kernels: ``"hip_resident"`` (the default) and ``"gpu_resident_batch"`` are
the resident solver, ``"baseline"`` and ``"fixed_dt_batch_barrier"`` the
reference's host-driven loops (``inundation_amd.hip.host_driven``), kept
for A/B checks. Open boundaries follow
``:74-121, 258-289, 361-435``: named stretches picked by a polyline, a
condition per stretch and phase, and the volume through each stretch at every
snapshot. Geometry comes from ``mesh_source`` (OBJ or (Geo)Parquet, through
the key-addressed ``geometry_cache_dir`` when set) or from a prebuilt
``geometry_cache_source`` artifact (``:658-760``); ``prepare()`` also takes
a built ``MeshGeometry`` directly.
"""

from __future__ import annotations

import os
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import TYPE_CHECKING, Any, Literal, cast

import numpy as np

from .mesh.boundaries import conveyance_split, select_boundary_edges
from .tuning import (
    DRY_TOL_DEFAULT,
    GRAVITY_G,
    MOMENTUM_SNAPSHOT_BACKENDS,
    OPEN_BOUNDARY_BACKENDS,
    OPEN_FREE,
    OPEN_INFLOW,
    OPEN_NORMAL,
    OPEN_STAGE,
    OPEN_WALL,
)

if TYPE_CHECKING:
    from collections.abc import Mapping, Sequence

    from numpy.typing import NDArray

    from .mesh.geometry import MeshGeometry

SOLVER_IMPLS = ("baseline", "fixed_dt_batch_barrier", "gpu_resident_batch", "hip_resident")
INITIAL_STATE_METADATA_KEYS = frozenset({"saved_order", "mesh_source", "t_end_s", "dt_max", "cfl_interval"})


@dataclass(frozen=True)
class PointSource:
    """Volumetric source over a circular area (discharge in m^3/s, centre in mesh coordinates)."""

    discharge_m3s: float
    center_xy: tuple[float, float]
    radius_m: float


@dataclass(frozen=True)
class OpenBoundary:
    """A stretch of mesh boundary that phases can open (reference ``workflow.py:74-86``).

    ``line`` is a polyline in the mesh file's coordinates; the boundary edges
    whose midpoints lie within ``tolerance_m`` of it form the stretch, which
    phases and results call ``name``.
    """

    name: str
    line: Sequence[tuple[float, float]]
    tolerance_m: float


@dataclass(frozen=True)
class Wall:
    """The stretch stays closed for this phase."""


@dataclass(frozen=True)
class FreeOutflow:
    """Zero-gradient outflow: exact for supercritical flow, partly reflecting for subcritical."""


@dataclass(frozen=True)
class NormalDepth:
    """Outflow at the Manning velocity of the local depth for an energy slope."""

    energy_slope: float


@dataclass(frozen=True)
class Stage:
    """A water level beyond the boundary [m, the bed's datum]."""

    elevation_m: float


@dataclass(frozen=True)
class Inflow:
    """Discharge into the mesh [m^3/s], spread over the stretch by Manning conveyance."""

    discharge_m3s: float
    energy_slope: float


BoundaryCondition = Wall | FreeOutflow | NormalDepth | Stage | Inflow


@dataclass(frozen=True)
class SimulationPhase:
    """One phase in a multi-phase hydrograph schedule."""

    duration_s: float
    sources: Sequence[PointSource]
    # The condition of every stretch of WorkflowConfig.open_boundaries, by name.
    boundaries: Mapping[str, BoundaryCondition] = field(default_factory=dict)


@dataclass(frozen=True)
class WorkflowConfig:
    """Runtime configuration for ``SWEWorkflow`` (reference ``workflow.py:135-227``)."""

    mesh_source: str | None
    manning_n: float
    output_interval_s: float
    dt_max: float
    cfl_interval: int
    gravity: float = GRAVITY_G
    dry_tol: float = DRY_TOL_DEFAULT
    cfl: float | None = None
    workgroup_size: int = 256
    dt_init: float = 1e-2
    use_hilbert_reorder: bool = True
    progress: bool = True
    area_tol: float = 1e-6
    initial_state_source: dict[str, object] | str | os.PathLike[str] | None = None
    initial_state_order: str = "solver"
    initial_state_id_col: str = "cell_id"
    track_clamp_mass: bool = False
    geometry_cache_dir: str | os.PathLike[str] | None = None
    geometry_cache_source: str | os.PathLike[str] | None = None
    capture_momentum_snapshots: bool = False
    open_boundaries: Sequence[OpenBoundary] = ()
    steps_per_graph: int | None = None
    deterministic: bool | None = None
    compensated: bool | None = None
    output_every_steps: int | None = None
    solver_impl: Literal["baseline", "fixed_dt_batch_barrier", "gpu_resident_batch", "hip_resident"] = (
        "hip_resident"
    )
    device: int = 0

    def __post_init__(self) -> None:
        if self.mesh_source is None and self.geometry_cache_source is None:
            raise ValueError(
                "mesh_source is required unless geometry_cache_source points at a "
                "prebuilt geometry-cache artifact"
            )
        if self.manning_n <= 0:
            raise ValueError(f"manning_n must be positive, got {self.manning_n}")
        if self.output_interval_s <= 0:
            raise ValueError(f"output_interval_s must be positive, got {self.output_interval_s}")
        if self.dt_max <= 0:
            raise ValueError(f"dt_max must be positive, got {self.dt_max}")
        if self.cfl_interval <= 0:
            raise ValueError(f"cfl_interval must be positive, got {self.cfl_interval}")
        if self.steps_per_graph is not None and self.steps_per_graph <= 0:
            raise ValueError(f"steps_per_graph must be positive, got {self.steps_per_graph}")
        if self.output_every_steps is not None and self.output_every_steps <= 0:
            raise ValueError(f"output_every_steps must be positive, got {self.output_every_steps}")
        if self.solver_impl not in SOLVER_IMPLS:
            raise ValueError(f"unsupported solver_impl: {self.solver_impl}")
        if self.capture_momentum_snapshots and self.solver_impl not in MOMENTUM_SNAPSHOT_BACKENDS:
            raise ValueError(
                f"capture_momentum_snapshots is not supported by solver_impl "
                f"{self.solver_impl!r}; use one of {sorted(MOMENTUM_SNAPSHOT_BACKENDS)}"
            )
        if self.open_boundaries:
            if self.solver_impl not in OPEN_BOUNDARY_BACKENDS:
                raise ValueError(
                    f"open_boundaries is not supported by solver_impl {self.solver_impl!r}; "
                    f"use one of {sorted(OPEN_BOUNDARY_BACKENDS)}"
                )
            if self.deterministic is False:
                raise ValueError("open_boundaries need deterministic accumulation")
            names = [b.name for b in self.open_boundaries]
            if len(set(names)) != len(names):
                raise ValueError(f"open boundary names repeat: {sorted(names)}")


@dataclass
class WorkflowResult:
    """Primary outputs (reference ``workflow.py:230-255``)."""

    snapshots: list[NDArray[np.float32]]
    snap_times: list[float]
    h_final: NDArray[np.float32]
    volume_final_m3: float
    volume_injected_m3: float
    wall_seconds: float
    steps_total: int
    hu_final: NDArray[np.float32] | None = None
    hv_final: NDArray[np.float32] | None = None
    hu_snapshots: list[NDArray[np.float32]] | None = None
    hv_snapshots: list[NDArray[np.float32]] | None = None
    clamp_mass_m3: float | None = None
    # Per open stretch, (snapshots, 2) cumulative volume [m^3] out of and into
    # the mesh through its edges, aligned with snap_times.
    boundary_volumes_m3: dict[str, NDArray[np.float64]] = field(default_factory=dict)


def inflow_phases(
    upstream: WorkflowResult,
    outlet: str,
    into: str,
    energy_slope: float,
    other_boundaries: Mapping[str, BoundaryCondition] | None = None,
    sources: Sequence[PointSource] = (),
) -> list[SimulationPhase]:
    """Phases that pass the volume leaving an upstream basin into the next one.

    Reference ``workflow.py:258-289``: one phase per output window of the
    upstream run, at the window's mean outflow through ``outlet``, entering
    through ``into``, so each window's volume passes whole.
    ``other_boundaries`` gives the downstream basin's other stretches.
    """
    t = np.asarray(upstream.snap_times, dtype=np.float64)
    out = upstream.boundary_volumes_m3[outlet][:, 0]
    # A phase's first snapshot repeats the previous phase's last one.
    ends = np.flatnonzero(np.diff(t) > 0.0) + 1
    starts = np.r_[0, ends[:-1]]
    phases = []
    for a, b in zip(starts.tolist(), ends.tolist(), strict=True):
        span = float(t[b] - t[a])
        cond = {**(other_boundaries or {}), into: Inflow(float((out[b] - out[a]) / span), energy_slope)}
        phases.append(SimulationPhase(span, sources, cond))
    return phases


def open_stretch_edges(geom: MeshGeometry, boundaries: Sequence[OpenBoundary]) -> list[NDArray[np.int64]]:
    """Each stretch's boundary edges (reference ``workflow.py:361-377``); empty or shared picks raise."""
    stretches = [select_boundary_edges(geom, b.line, b.tolerance_m) for b in boundaries]
    for b, edges in zip(boundaries, stretches, strict=True):
        if edges.size == 0:
            raise ValueError(
                f"open boundary {b.name!r} selects no boundary edges within {b.tolerance_m} m of its line"
            )
    picked = np.concatenate(stretches) if stretches else np.empty(0, dtype=np.int64)
    if np.unique(picked).size != picked.size:
        raise ValueError("open boundaries select the same edge twice")
    return stretches


def open_conditions(
    geom: MeshGeometry,
    manning: NDArray[np.float32],
    boundary: OpenBoundary,
    condition: BoundaryCondition,
    edges: NDArray[np.int64],
) -> tuple[int, NDArray[np.float32]]:
    """Condition code and per-edge value of one stretch (reference ``workflow.py:380-410``).

    The value is what the ghost state reads: sqrt(S)/n for normal depth, the
    depth of the stage over the edge cell's bed (at least 0), the inflow's
    discharge per unit width from ``conveyance_split``.
    """
    cells = np.asarray(geom.edge_cellL)[edges]
    fields = [float(getattr(condition, f)) for f in getattr(condition, "__dataclass_fields__", {})]
    if not np.isfinite(fields).all():
        raise ValueError(f"open boundary {boundary.name!r} has a non-finite value: {condition}")
    if isinstance(condition, Wall):
        return OPEN_WALL, np.zeros(edges.size, dtype=np.float32)
    if isinstance(condition, FreeOutflow):
        return OPEN_FREE, np.zeros(edges.size, dtype=np.float32)
    if isinstance(condition, NormalDepth):
        if condition.energy_slope <= 0.0:
            raise ValueError(f"open boundary {boundary.name!r} needs a positive energy slope")
        return OPEN_NORMAL, (np.sqrt(condition.energy_slope) / manning[cells]).astype(np.float32)
    if isinstance(condition, Stage):
        depth = condition.elevation_m - np.asarray(geom.zb)[cells].astype(np.float64)
        return OPEN_STAGE, np.maximum(depth, 0.0).astype(np.float32)
    if not isinstance(condition, Inflow):
        raise TypeError(f"open boundary {boundary.name!r}: unknown condition {condition!r}")
    if condition.energy_slope <= 0.0:
        raise ValueError(f"open boundary {boundary.name!r} needs a positive energy slope")
    if condition.discharge_m3s < 0.0:
        raise ValueError(f"open boundary {boundary.name!r} has a negative inflow")
    return OPEN_INFLOW, conveyance_split(
        condition.discharge_m3s,
        np.asarray(geom.edge_len)[edges],
        np.asarray(geom.zb)[cells],
        manning[cells],
        condition.energy_slope,
    )


def phase_open_conditions(
    geom: MeshGeometry,
    manning: NDArray[np.float32],
    boundaries: Sequence[OpenBoundary],
    stretches: Sequence[NDArray[np.int64]],
    phase: SimulationPhase,
) -> tuple[NDArray[np.uint32], NDArray[np.float32]]:
    """Per-edge condition codes and values of one phase, in stretch order (reference ``:413-435``)."""
    names = {b.name for b in boundaries}
    if set(phase.boundaries) != names:
        raise ValueError(
            f"phase boundaries {sorted(phase.boundaries)} must name every open boundary: {sorted(names)}"
        )
    kinds, values = [], []
    for b, edges in zip(boundaries, stretches, strict=True):
        kind, value = open_conditions(geom, manning, b, phase.boundaries[b.name], edges)
        kinds.append(np.full(edges.size, kind, dtype=np.uint32))
        values.append(value)
    return np.concatenate(kinds), np.concatenate(values)


def _point_in_polygon(point: NDArray[np.float64], polygon: NDArray[np.float64]) -> bool:
    """Inside or on the boundary of a polygon, even-odd rule (reference ``workflow.py:294-314``)."""
    x, y = float(point[0]), float(point[1])
    scale = max(1.0, float(np.abs(polygon).max()), abs(x), abs(y))
    tol = 1e-10 * scale
    inside = False
    for a, b in zip(polygon, np.roll(polygon, -1, axis=0)):
        ax, ay, bx, by = float(a[0]), float(a[1]), float(b[0]), float(b[1])
        cross = (x - ax) * (by - ay) - (y - ay) * (bx - ax)
        if (
            abs(cross) <= tol
            and min(ax, bx) - tol <= x <= max(ax, bx) + tol
            and min(ay, by) - tol <= y <= max(ay, by) + tol
        ):
            return True
        if (ay > y) != (by > y) and x < ax + (y - ay) * (bx - ax) / (by - ay):
            inside = not inside
    return inside


def select_cells_within_radius(
    centroid: NDArray[np.float32],
    center_xy: tuple[float, float],
    radius_m: float,
    valid_mask: NDArray[np.bool_] | None = None,
    cell_vertices: NDArray[np.float32] | None = None,
    cell_vertex_ptr: NDArray[np.int32] | None = None,
    cell_bbox: NDArray[np.float32] | None = None,
) -> NDArray[np.bool_]:
    """Cells whose centroid lies within the radius, plus the cell containing the centre.

    Reference ``workflow.py:317-358``.
    """
    center = np.asarray(center_xy, dtype=np.float64)
    if radius_m < 0.0:
        return np.zeros(centroid.shape[0], dtype=bool)
    mask = np.hypot(centroid[:, 0] - center[0], centroid[:, 1] - center[1]) <= float(radius_m)
    if valid_mask is not None:
        mask &= valid_mask
    if cell_vertices is None or cell_vertex_ptr is None or cell_bbox is None or not centroid.shape[0]:
        return mask
    bb = np.asarray(cell_bbox, dtype=np.float64)
    cand = (bb[:, 0] <= center[0]) & (bb[:, 1] >= center[0]) & (bb[:, 2] <= center[1]) & (bb[:, 3] >= center[1])
    if valid_mask is not None:
        cand &= valid_mask
    for c in np.flatnonzero(cand):
        poly = np.asarray(cell_vertices[int(cell_vertex_ptr[c]) : int(cell_vertex_ptr[c + 1])], dtype=np.float64)
        if _point_in_polygon(center, poly):
            mask[c] = True
    return mask


def build_phase_source_rate(
    geom: MeshGeometry, phase: SimulationPhase, valid_cell: NDArray[np.bool_]
) -> NDArray[np.float32]:
    """Per-cell source [m^3/s] of one phase, split by cell area (reference ``workflow.py:599-625``)."""
    src = np.zeros(geom.N, dtype=np.float32)
    ox, oy = geom.origin.tolist()
    for s in phase.sources:
        cx, cy = s.center_xy
        mask = select_cells_within_radius(
            geom.centroid,
            (cx - ox, cy - oy),
            s.radius_m,
            valid_mask=valid_cell,
            cell_vertices=geom.cell_vertices,
            cell_vertex_ptr=geom.cell_vertex_ptr,
            cell_bbox=geom.cell_bbox,
        )
        if not mask.any():
            continue
        a = geom.area[mask]
        total = float(a.sum())
        if total <= 0.0:
            continue
        src[mask] += np.float32(s.discharge_m3s) * (a / total)
    return src


def _as_1d_float32(name: str, values: object, n: int) -> NDArray[np.float32]:
    arr = np.asarray(values, dtype=np.float32).reshape(-1)
    if arr.size != n:
        raise ValueError(f"{name}: expected {n} values, got {arr.size}")
    if not np.isfinite(arr).all():
        raise ValueError(f"{name}: contains non-finite values")
    return arr


def load_initial_state(
    source: dict[str, object] | str | os.PathLike[str] | None,
    n_cells: int,
    zb: NDArray[np.float32],
    *,
    order: str,
    id_col: str,
    perm: NDArray[np.int32],
    dry_tol: float,
) -> tuple[NDArray[np.float32], NDArray[np.float32], NDArray[np.float32], NDArray[np.float32] | None] | None:
    """Optional per-cell initial state (h or wse, hu/hv or u/v, optional n_mann).

    Reference ``workflow.py:450-596``: a dict, ``.npz``, ``.npy`` ((N, 3) or
    (N, 4) columns), ``.csv`` or ``.parquet`` (through pandas). Metadata
    entries (``saved_order``, ``mesh_source``, ``t_end_s``, ``dt_max``,
    ``cfl_interval``) are ignored; an ``id_col`` column maps rows to cells;
    ``order="original"`` takes rows in the mesh file's cell order.
    """
    if source is None:
        return None
    if isinstance(source, dict):
        state = {k: np.asarray(v) for k, v in source.items()}
    else:
        path = Path(os.fspath(source))
        if not path.exists():
            raise FileNotFoundError(f"Initial state file not found: {path}")
        suffix = path.suffix.lower()
        if suffix == ".npz":
            with np.load(path) as npz:
                state = {k: npz[k] for k in npz.files}
        elif suffix == ".npy":
            arr = np.load(path)
            if arr.ndim != 2 or arr.shape[1] not in (3, 4):
                raise ValueError("For .npy input, expected shape (N,3)=[h,hu,hv] or (N,4)=[h,hu,hv,n_mann]")
            state = {"h": arr[:, 0], "hu": arr[:, 1], "hv": arr[:, 2]}
            if arr.shape[1] == 4:
                state["n_mann"] = arr[:, 3]
        elif suffix in (".csv", ".parquet"):
            try:
                import pandas as pd
            except ImportError as exc:
                raise RuntimeError("pandas is required to load .csv/.parquet initial state files") from exc
            df = pd.read_csv(path) if suffix == ".csv" else pd.read_parquet(path)
            state = {c: df[c].to_numpy() for c in df.columns}
        else:
            raise ValueError(f"Unsupported initial state format: {suffix}. Use .npz, .npy, .csv, or .parquet")
    valid = {"h", "wse", "hu", "hv", "u", "v", "n_mann", id_col}
    state = {k: v for k, v in state.items() if k not in INITIAL_STATE_METADATA_KEYS}
    bad = sorted(k for k in state if k not in valid)
    if bad:
        raise ValueError(f"initial state contains unsupported fields: {bad}. Supported fields: {sorted(valid)}")
    if id_col in state:
        ids_f = np.asarray(state[id_col]).reshape(-1).astype(np.float64)
        if not np.isfinite(ids_f).all() or not np.equal(ids_f, np.floor(ids_f)).all():
            raise ValueError(f"{id_col}: values must be finite integers")
        ids = ids_f.astype(np.int64)
        if ids.size != n_cells or np.unique(ids).size != n_cells or ids.min() < 0 or ids.max() >= n_cells:
            raise ValueError(f"{id_col}: must be a permutation of [0, {n_cells - 1}]")
        aligned = {}
        for k, v in state.items():
            if k == id_col:
                continue
            src = np.asarray(v).reshape(-1)
            if src.size != n_cells:
                raise ValueError(f"{k}: expected {n_cells} rows, got {src.size}")
            dst = np.empty(n_cells, dtype=src.dtype)
            dst[ids] = src
            aligned[k] = dst
        state = aligned
    if "h" in state:
        h = _as_1d_float32("h", state["h"], n_cells)
    elif "wse" in state:
        h = _as_1d_float32("wse", state["wse"], n_cells) - zb.astype(np.float32)
    else:
        raise ValueError("Initial state must provide 'h' or 'wse'")
    if "hu" in state and "hv" in state:
        hu = _as_1d_float32("hu", state["hu"], n_cells)
        hv = _as_1d_float32("hv", state["hv"], n_cells)
    elif "u" in state and "v" in state:
        hu = h * _as_1d_float32("u", state["u"], n_cells)
        hv = h * _as_1d_float32("v", state["v"], n_cells)
    else:
        raise ValueError("Initial state must provide (hu,hv) or (u,v)")
    n_mann = _as_1d_float32("n_mann", state["n_mann"], n_cells) if "n_mann" in state else None
    if order not in ("solver", "original"):
        raise ValueError("initial_state_order must be 'solver' or 'original'")
    if order == "original":
        h, hu, hv = h[perm], hu[perm], hv[perm]
        if n_mann is not None:
            n_mann = n_mann[perm]
    h = np.maximum(h, 0.0).astype(np.float32)
    wet = h > np.float32(dry_tol)
    hu = np.where(wet, hu, 0.0).astype(np.float32)
    hv = np.where(wet, hv, 0.0).astype(np.float32)
    if n_mann is not None:
        n_mann = np.maximum(n_mann.astype(np.float32), np.float32(1e-4))
    return h, hu, hv, n_mann


def save_warm_start(
    path: str | os.PathLike[str],
    result: WorkflowResult,
    perm: NDArray[np.int32],
    *,
    order: str = "original",
    **metadata: object,
) -> Path:
    """A run's final state as an ``.npz`` warm start that ``initial_state_source`` loads.

    ``order="original"`` writes rows in the mesh file's cell order (load it
    with ``initial_state_order="original"``), ``"solver"`` in the solver's.
    ``saved_order`` and any ``metadata`` (for example ``mesh_source``,
    ``t_end_s``) ride along and are ignored by the loader.
    """
    if result.hu_final is None or result.hv_final is None:
        raise ValueError("the result holds no final momentum")
    if order not in ("solver", "original"):
        raise ValueError("order must be 'solver' or 'original'")
    fields = {"h": result.h_final, "hu": result.hu_final, "hv": result.hv_final}
    if order == "original":
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size, dtype=perm.dtype)
        fields = {k: v[inv] for k, v in fields.items()}
    bad = sorted(set(metadata) - INITIAL_STATE_METADATA_KEYS)
    if bad:
        raise ValueError(f"unknown warm-start metadata {bad}; allowed: {sorted(INITIAL_STATE_METADATA_KEYS)}")
    out = Path(os.fspath(path))
    np.savez(out, saved_order=np.asarray(order), **{k: np.asarray(v) for k, v in metadata.items()}, **fields)
    return out


def load_geometry(cfg: WorkflowConfig) -> tuple[MeshGeometry, NDArray[np.int32]]:
    """The solver's geometry and ``perm[new] = old`` (reference ``workflow.py:658-752``).

    A ``geometry_cache_source`` artifact wins and the mesh file is never
    opened; its cell order must match ``use_hilbert_reorder``. Otherwise the
    mesh file is loaded, built and reordered, through the directory cache
    when ``geometry_cache_dir`` is set (a hit skips the build, a miss stores
    the result).
    """
    from .mesh.cache import (
        geometry_cache_key,
        geometry_from_mesh_file,
        load_geometry_cache,
        load_geometry_cache_artifact,
        save_geometry_cache,
    )

    def log(msg: str) -> None:
        if cfg.progress:
            sys.stdout.write(f"[prepare] {msg}\n")
            sys.stdout.flush()

    if cfg.geometry_cache_source is not None:
        geom, perm, meta = load_geometry_cache_artifact(cfg.geometry_cache_source)
        if meta.use_hilbert_reorder != cfg.use_hilbert_reorder:
            raise ValueError(
                f"geometry cache artifact {cfg.geometry_cache_source} was built with "
                f"use_hilbert_reorder={meta.use_hilbert_reorder}, but WorkflowConfig.use_hilbert_reorder "
                f"is {cfg.use_hilbert_reorder} — rebuild the artifact or match the config"
            )
        log(f"geometry cache artifact loaded ({cfg.geometry_cache_source}; built from {meta.mesh_source})")
        return geom, perm
    if cfg.mesh_source is None:
        raise RuntimeError("mesh_source is None without a geometry cache source")
    key = ""
    if cfg.geometry_cache_dir is not None:
        key = geometry_cache_key(cfg.mesh_source, use_hilbert_reorder=cfg.use_hilbert_reorder)
        hit = load_geometry_cache(cfg.geometry_cache_dir, key)
        if hit is not None:
            log(f"geometry cache hit ({cfg.geometry_cache_dir}, key={key[:12]}...)")
            return hit
        log(f"geometry cache miss ({cfg.geometry_cache_dir}, key={key[:12]}...) — building")
    t0 = time.perf_counter()
    geom, perm = geometry_from_mesh_file(cfg.mesh_source, use_hilbert_reorder=cfg.use_hilbert_reorder)
    log(f"geometry built in {time.perf_counter() - t0:.1f}s ({cfg.mesh_source})")
    if cfg.geometry_cache_dir is not None:
        path = save_geometry_cache(cfg.geometry_cache_dir, key, geom, perm)
        log(f"geometry cache saved ({path})")
    return geom, perm


class SWEWorkflow:
    """Prepare once, run phase schedules (reference ``workflow.py:628-969``)."""

    def __init__(self, config: WorkflowConfig) -> None:
        self.config = config
        self.geom: MeshGeometry | None = None
        self.perm: NDArray[np.int32] | None = None
        self.solver: Any = None
        self._open_stretches: list[NDArray[np.int64]] = []
        self._manning: NDArray[np.float32] | None = None

    def prepare(self, geometry: MeshGeometry | None = None, perm: NDArray[np.int32] | None = None) -> None:
        """Geometry (given, or loaded as ``load_geometry`` does), initial state, solver."""
        cfg = self.config
        if geometry is None:
            geometry, perm = load_geometry(cfg)
        geom = geometry
        if perm is None:
            perm = np.arange(geom.N, dtype=np.int32)
        h0 = np.zeros(geom.N, dtype=np.float32)
        hu0 = np.zeros(geom.N, dtype=np.float32)
        hv0 = np.zeros(geom.N, dtype=np.float32)
        n0 = geom.manning_n if geom.manning_n is not None else np.full(geom.N, cfg.manning_n, dtype=np.float32)
        loaded = load_initial_state(
            cfg.initial_state_source,
            geom.N,
            geom.zb,
            order=cfg.initial_state_order,
            id_col=cfg.initial_state_id_col,
            perm=perm,
            dry_tol=cfg.dry_tol,
        )
        if loaded is not None:
            h0, hu0, hv0, n_loaded = loaded
            if n_loaded is not None:
                n0 = n_loaded
        from .hip.host_driven import SWESolverHipHostDriven
        from .hip.solver import SWESolverHipResident

        self._open_stretches = open_stretch_edges(geom, cfg.open_boundaries)
        self._manning = np.asarray(n0, dtype=np.float32)
        args = (geom, None, h0, hu0, hv0, n0)
        kwargs: dict[str, Any] = {
            "g": cfg.gravity,
            "dry_tol": cfg.dry_tol,
            "cfl": cfg.cfl,
            "workgroup_size": cfg.workgroup_size,
            "open_edges": self._open_stretches,
            "device": cfg.device,
        }
        resident = cfg.solver_impl in ("hip_resident", "gpu_resident_batch")
        if resident:
            solver: Any = SWESolverHipResident(*args, **kwargs)
        else:
            solver = SWESolverHipHostDriven(*args, mode=cfg.solver_impl, **kwargs)
        if cfg.track_clamp_mass:
            solver.track_clamp = True
        # The resident backend takes the run-recording attributes
        # (reference workflow.py:816-826).
        if resident:
            if cfg.steps_per_graph is not None:
                solver.steps_per_graph = cfg.steps_per_graph
            if cfg.deterministic is not None:
                solver.deterministic = cfg.deterministic
            if cfg.compensated is not None:
                solver.compensated = cfg.compensated
        self.geom = geom
        self.perm = perm
        self.solver = solver

    def run(self, phases: Sequence[SimulationPhase]) -> WorkflowResult:
        """Execute the phases in order and return the outputs (reference ``workflow.py:828-969``)."""
        if self.geom is None or self.solver is None:
            raise RuntimeError("Workflow not prepared. Call prepare() first.")
        if not phases:
            raise ValueError("At least one phase is required.")
        cfg = self.config
        area = self.geom.area
        valid = area > cfg.area_tol
        snaps: list[NDArray[np.float32]] = []
        times: list[float] = []
        hu_s: list[NDArray[np.float32]] | None = [] if cfg.capture_momentum_snapshots else None
        hv_s: list[NDArray[np.float32]] | None = [] if cfg.capture_momentum_snapshots else None
        t = 0.0
        wall0 = time.perf_counter()
        for k, phase in enumerate(phases):
            if phase.duration_s <= 0.0:
                raise ValueError(f"Phase {k} has non-positive duration.")
            t_end = t + phase.duration_s
            src = build_phase_source_rate(self.geom, phase, valid)
            if self._open_stretches:
                self.solver.set_open_conditions(
                    *phase_open_conditions(
                        self.geom, self._manning, cfg.open_boundaries, self._open_stretches, phase
                    )
                )
            res = self.solver.run(
                t_end=t_end,
                output_interval_s=cfg.output_interval_s,
                dt_max=cfg.dt_max,
                dt_init=cfg.dt_init,
                progress=cfg.progress,
                source_rate=src,
                resume=k > 0,
                t_start=t,
                cfl_interval=cfg.cfl_interval,
                output_every_steps=cfg.output_every_steps if cfg.solver_impl == "fixed_dt_batch_barrier" else None,
                capture_momentum_snapshots=cfg.capture_momentum_snapshots,
            )
            if len(res) == 4:
                ps, pt, phu, phv = res
                if hu_s is not None and hv_s is not None:
                    hu_s.extend(phu)
                    hv_s.extend(phv)
            else:
                ps, pt = res
            snaps.extend(ps)
            times.extend(pt)
            t = t_end
        wall = time.perf_counter() - wall0
        h_final = self.solver.download_h()
        h_precise = cast("NDArray[np.float64]", self.solver.download_h_precise())
        vol = float((h_precise * area.astype(np.float64)).sum())
        injected = float(sum(p.duration_s * sum(s.discharge_m3s for s in p.sources) for p in phases))
        clamp = None
        if cfg.track_clamp_mass:
            clamp = float(np.asarray(self.solver.download_clamp_mass(), dtype=np.float64).sum())
        volumes: dict[str, NDArray[np.float64]] = {}
        if self._open_stretches:
            stacked = np.stack(self.solver.open_volume_snapshots)
            volumes = {b.name: stacked[:, k] for k, b in enumerate(cfg.open_boundaries)}
        if cfg.progress:
            sys.stdout.write(f"[workflow] {self.solver.steps_total} steps in {wall:.2f}s\n")
        return WorkflowResult(
            snapshots=snaps,
            snap_times=times,
            h_final=h_final,
            volume_final_m3=vol,
            volume_injected_m3=injected,
            wall_seconds=wall,
            steps_total=int(self.solver.steps_total),
            hu_final=self.solver.download_hu(),
            hv_final=self.solver.download_hv(),
            hu_snapshots=hu_s,
            hv_snapshots=hv_s,
            clamp_mass_m3=clamp,
            boundary_volumes_m3=volumes,
        )


__all__ = [
    "BoundaryCondition",
    "FreeOutflow",
    "Inflow",
    "NormalDepth",
    "OpenBoundary",
    "PointSource",
    "Stage",
    "Wall",
    "inflow_phases",
    "load_geometry",
    "open_conditions",
    "open_stretch_edges",
    "phase_open_conditions",
    "save_warm_start",
    "SWEWorkflow",
    "SimulationPhase",
    "WorkflowConfig",
    "WorkflowResult",
    "build_phase_source_rate",
    "load_initial_state",
    "select_cells_within_radius",
]
