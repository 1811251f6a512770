"""``hip_resident``: the GPU-resident Heun solver on MI355X through the C ABI.

Same backend protocol as the reference's ``SWESolverGpuResidentBatch``
(``inundation/vulkan/gpu_resident_batch.py:140-762``): the constructor takes
the geometry and initial state, the workflow sets ``track_clamp``,
``steps_per_graph``, ``deterministic`` and ``compensated`` afterwards, and
``run()`` advances output window by output window with the on-device integer
clock, replaying a hipGraph of ``steps_per_graph`` steps until each window's
stop tick is reached. Steps recorded past a window's stop have dt = 0 and do
nothing, exactly as in the reference.
"""

from __future__ import annotations


import sys
import time as _time
from typing import TYPE_CHECKING, Any

import numpy as np

from ..mesh.blocks import edge_order_by_block, num_blocks, remap_cell_edges
from ..tuning import (
    ACTIVE_BLOCK_CELLS,
    CFL_DEFAULT,
    DRY_TOL_DEFAULT,
    GRAVITY_G,
    dt_to_ticks,
    output_window_stops,
    seconds_to_ticks,
    ticks_to_seconds,
)
from . import lib as L

if TYPE_CHECKING:
    from collections.abc import Callable, Sequence

    from numpy.typing import NDArray

    from ..mesh.geometry import MeshGeometry

STEPS_PER_GRAPH_DEFAULT = 200


def device_mesh_arrays(geom: MeshGeometry) -> dict[str, Any]:
    """The mesh in the device's edge order (stable by the left cell's block).

    Reference ``gpu_resident_batch.py:173-192`` and ``:226-229``: edges sorted
    by ``edge_order_by_block``, CSR entries renumbered by ``remap_cell_edges``
    with each cell's slot order kept.
    """
    order = edge_order_by_block(geom.edge_cellL, ACTIVE_BLOCK_CELLS)
    return {
        "order": order,
        "cell_l": np.ascontiguousarray(geom.edge_cellL[order], dtype=np.int32),
        "cell_r": np.ascontiguousarray(geom.edge_cellR[order], dtype=np.int32),
        "nx": np.ascontiguousarray(geom.edge_nx[order], dtype=np.float32),
        "ny": np.ascontiguousarray(geom.edge_ny[order], dtype=np.float32),
        "len": np.ascontiguousarray(geom.edge_len[order], dtype=np.float32),
        "dz2": np.ascontiguousarray(
            np.column_stack([geom.edge_dzL[order], geom.edge_dzR[order]]).ravel(), dtype=np.float32
        ),
        "area": np.ascontiguousarray(geom.area, dtype=np.float32),
        "cptr": np.ascontiguousarray(geom.cell_edge_ptr, dtype=np.uint32),
        "cidx": remap_cell_edges(geom.cell_edge_idx_signed, order),
    }


def mark_open_edges(
    geom: MeshGeometry, m: dict[str, Any], open_edges: Sequence[NDArray[np.int64]]
) -> NDArray[np.uint32]:
    """Open edge k as ``cell_r == -2 - k`` in ``m`` (the device edge order), stretches in order.

    Reference ``vulkan/solver.py:165, 238-243`` and ``gpu_resident_batch.py:176-181``:
    ``open_edges`` holds each stretch's edges as indices into ``geom``; every
    other boundary edge stays a wall. Returns the stretch pointer (open edges
    ``[ptr[s], ptr[s + 1])`` form stretch s).
    """
    stretches = [np.asarray(e, dtype=np.int64).ravel() for e in open_edges]
    edges = np.concatenate([np.empty(0, dtype=np.int64), *stretches])
    if edges.size:
        if edges.min() < 0 or edges.max() >= geom.E:
            raise ValueError("open edge index out of range")
        if np.unique(edges).size != edges.size:
            raise ValueError("an edge is in more than one open stretch")
        if np.any(np.asarray(geom.edge_cellR)[edges] != -1):
            raise ValueError("open edges must be boundary edges")
        if any(st.size == 0 for st in stretches):
            raise ValueError("every open stretch needs an edge")
        position = np.empty(geom.E, dtype=np.int64)
        position[m["order"]] = np.arange(geom.E)
        cell_r = m["cell_r"].copy()
        cell_r[position[edges]] = (-2 - np.arange(edges.size)).astype(np.int32)
        m["cell_r"] = cell_r
    return np.r_[0, np.cumsum([st.size for st in stretches], dtype=np.int64)].astype(np.uint32)


class SWESolverHipResident:
    """Record once, replay many steps; dt and time stay on the GPU."""

    default_cfl: float = CFL_DEFAULT
    steps_per_graph: int = STEPS_PER_GRAPH_DEFAULT
    active_set: bool = True
    deterministic: bool = True
    compensated: bool = True
    cfl_safety: float = 1.0

    def __init__(
        self,
        geom: MeshGeometry,
        spv: dict[str, bytes] | None,
        h0: NDArray[np.float32],
        hu0: NDArray[np.float32],
        hv0: NDArray[np.float32],
        n0: NDArray[np.float32],
        *,
        g: float = GRAVITY_G,
        dry_tol: float = DRY_TOL_DEFAULT,
        cfl: float | None = None,
        workgroup_size: int = 256,
        open_edges: Sequence[NDArray[np.int64]] = (),
        device: int = 0,
    ) -> None:
        del spv  # kernels are compiled ahead of time
        if workgroup_size != ACTIVE_BLOCK_CELLS:
            raise ValueError(
                f"hip_resident needs workgroup_size={ACTIVE_BLOCK_CELLS}, got {workgroup_size}"
            )
        self.geom = geom
        self.N = geom.N
        self.E = geom.E
        self.n_blocks = num_blocks(self.N, ACTIVE_BLOCK_CELLS)
        self._g = float(g)
        self._dry_tol = float(dry_tol)
        self._cfl = float(self.default_cfl if cfl is None else cfl)
        self.track_clamp = False
        self.open_volume_snapshots: list[NDArray[np.float64]] = []
        self._area = geom.area
        self._h0 = np.asarray(h0, dtype=np.float32).ravel().copy()
        self._hu0 = np.asarray(hu0, dtype=np.float32).ravel().copy()
        self._hv0 = np.asarray(hv0, dtype=np.float32).ravel().copy()
        m = device_mesh_arrays(geom)
        stretch_ptr = mark_open_edges(geom, m, open_edges)
        self._open_stretches = [np.asarray(e, dtype=np.int64).ravel() for e in open_edges]
        self.n_open = int(stretch_ptr[-1])
        self._open_cells = np.asarray(geom.edge_cellL)[
            np.concatenate([np.empty(0, dtype=np.int64), *self._open_stretches])
        ]
        self._ctx = self._open_context(device, m)
        if self.n_open:
            self._ctx.set_open_stretches(stretch_ptr)
        self._ctx.set_manning(np.asarray(n0, dtype=np.float32))
        self._apply_params()
        self._ctx.set_state(self._h0, self._hu0, self._hv0, None)
        self._eft_checked = False
        self.steps_total = 0

    def _open_context(self, device: int, m: dict[str, Any]) -> Any:
        """A device context on ``device`` holding mesh ``m`` (the device edge order, open edges marked)."""
        ctx = L.Context(device)
        ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
        return ctx

    # -- parameters ---------------------------------------------------------

    def _flags(self) -> int:
        """Kernel flags; ``deterministic = False`` sums edge words in arrival order (not bit-reproducible)."""
        if self.n_open and not self.deterministic:
            raise ValueError("open boundaries need deterministic accumulation")
        f = L.FLAG_DETERMINISTIC if self.deterministic else 0
        if self.compensated:
            f |= L.FLAG_COMPENSATED
        if self.active_set:
            f |= L.FLAG_ACTIVE_SET
        if self.track_clamp:
            f |= L.FLAG_TRACK_CLAMP
        return f

    def _apply_params(self) -> None:
        self._ctx.set_params(self._g, self._dry_tol, self._cfl, self._flags())

    # -- state --------------------------------------------------------------

    def reset(self, h0: NDArray[np.float32], hu0: NDArray[np.float32], hv0: NDArray[np.float32]) -> None:
        """Re-upload initial conditions; the compensated low word starts at zero."""
        self._h0 = np.asarray(h0, dtype=np.float32).ravel().copy()
        self._hu0 = np.asarray(hu0, dtype=np.float32).ravel().copy()
        self._hv0 = np.asarray(hv0, dtype=np.float32).ravel().copy()
        self._ctx.set_state(self._h0, self._hu0, self._hv0, None)

    def _reset_to_initial(self) -> None:
        self.reset(self._h0, self._hu0, self._hv0)

    def set_open_conditions(self, kinds: NDArray[np.uint32], values: NDArray[np.float32]) -> None:
        """Every open edge's condition (``L.OPEN_*``) and value, in stretch order, before a run.

        Reference ``vulkan/solver.py:427-437`` and ``gpu_resident_batch.py:447-458``:
        a block with an edge that can bring water in stays active while these hold.
        """
        kinds = np.asarray(kinds, dtype=np.uint32).ravel()
        values = np.asarray(values, dtype=np.float32).ravel()
        if kinds.size != self.n_open or values.size != self.n_open:
            raise ValueError(
                f"expected {self.n_open} open edges, got {kinds.size} kinds and {values.size} values"
            )
        if self.n_open:
            self._ctx.set_open_conditions(kinds, values)

    def download_open_volumes(self) -> NDArray[np.float64]:
        """``(stretches, 2)`` volume [m^3] out of and into the mesh since the start."""
        return self._ctx.open_volumes(len(self._open_stretches))

    def _check_error_free_arithmetic(self) -> None:
        """Refuse to run if two-sum lost its error term (reference solver.py:449-467)."""
        if self._eft_checked:
            return
        out = self._ctx.probe(np.array([[1.0, 1e-9]], dtype=np.float32), self._g, self._dry_tol)
        if float(out[0, 0]) != float(np.float32(1e-9)):
            raise RuntimeError(
                "error-free arithmetic is broken on this device, so compensated depth would "
                "leak mass: set compensated = False"
            )
        self._eft_checked = True

    # -- downloads ----------------------------------------------------------

    def download_h(self) -> NDArray[np.float32]:
        return self._ctx.download(L.FIELD_H, self.N)

    def download_hu(self) -> NDArray[np.float32]:
        return self._ctx.download(L.FIELD_HU, self.N)

    def download_hv(self) -> NDArray[np.float32]:
        return self._ctx.download(L.FIELD_HV, self.N)

    def download_h_residual(self) -> NDArray[np.float32]:
        """Low words of the compensated depth; zero unless ``compensated`` ran."""
        return self._ctx.download(L.FIELD_H_LO, self.N)

    def download_h_precise(self) -> NDArray[np.float64]:
        """Depth as float64: the committed word plus its low word."""
        return self.download_h().astype(np.float64) + self.download_h_residual().astype(np.float64)

    def download_block_flags(self) -> NDArray[np.uint32]:
        """``(n_blocks, 3)`` active-set flags: wet, active, forced."""
        cols = [
            self._ctx.download(w, self.n_blocks, np.uint32)
            for w in (L.FIELD_BLOCK_WET, L.FIELD_BLOCK_ACTIVE, L.FIELD_BLOCK_FORCED)
        ]
        return np.column_stack(cols)

    def download_clamp_mass(self) -> NDArray[np.float32]:
        return self._ctx.download(L.FIELD_CLAMP, self.N)

    def reset_clamp_mass(self) -> None:
        self._ctx.upload(L.FIELD_CLAMP, np.zeros(self.N, dtype=np.float32))

    def read_ctrl(self) -> L.InuCtrl:
        return self._ctx.ctrl_read()

    @property
    def context(self) -> L.Context:
        return self._ctx

    # -- run ----------------------------------------------------------------

    def _set_sources(self, source_rate: NDArray[np.float32] | None) -> int:
        """(cell, dh/dt) pairs of the nonzero sources (gpu_resident_batch.py:529-560, 665-670)."""
        if source_rate is None:
            self._ctx.set_sources(np.empty(0, np.uint32), np.empty(0, np.float32))
            return 0
        per_area = (1.0 / self._area).astype(np.float32)
        dh = np.asarray(source_rate, dtype=np.float32) * per_area
        cells = np.flatnonzero(dh).astype(np.uint32)
        self._ctx.set_sources(cells, dh[cells].astype(np.float32))
        return int(cells.size)

    def run(
        self,
        t_end: float,
        output_interval_s: float,
        dt_max: float,
        dt_init: float,
        cfl_interval: int,
        progress: bool = True,
        source_rate: NDArray[np.float32] | None = None,
        source_fn: Callable[[float], NDArray[np.float32]] | None = None,
        resume: bool = False,
        t_start: float = 0.0,
        capture_momentum_snapshots: bool = False,
        output_every_steps: int | None = None,
    ) -> Any:
        """Advance to ``t_end``; snapshot every ``output_interval_s`` from the run's start.

        Reference ``gpu_resident_batch.py:609-762``. ``cfl_interval`` and
        ``output_every_steps`` are accepted for interface compatibility and
        ignored. Returns ``(snapshots, times)`` or, with
        ``capture_momentum_snapshots``, ``(snapshots, times, hu, hv)``.
        """
        del cfl_interval, output_every_steps
        if source_fn is not None:
            raise ValueError("hip_resident takes a constant source_rate per run, not source_fn")
        if self.steps_per_graph < 1:
            raise ValueError(f"steps_per_graph must be positive, got {self.steps_per_graph}")
        dt_max_ticks = dt_to_ticks(dt_max)
        dry_dt_ticks = dt_to_ticks(min(dt_init, dt_max))
        start_ticks = seconds_to_ticks(t_start if resume else 0.0)
        end_ticks = seconds_to_ticks(t_end)
        interval_ticks = seconds_to_ticks(output_interval_s)

        self._apply_params()
        if not resume:
            self._reset_to_initial()
            self.open_volume_snapshots = []
        if self.compensated or self.n_open:
            self._check_error_free_arithmetic()
        self._set_sources(source_rate)

        snapshots = [self.download_h() if resume else self._h0.copy()]
        snap_times = [ticks_to_seconds(start_ticks)]
        hu_snaps: list[NDArray[np.float32]] = []
        hv_snaps: list[NDArray[np.float32]] = []

        def capture() -> None:
            if capture_momentum_snapshots:
                hu_snaps.append(self.download_hu())
                hv_snaps.append(self.download_hv())
            if self.n_open:
                self.open_volume_snapshots.append(self.download_open_volumes())

        capture()
        ctrl = L.InuCtrl()
        ctrl.now_ticks = start_ticks
        ctrl.steps = 0
        ctrl.dt_ticks = 0
        ctrl.dt_max_ticks = dt_max_ticks
        ctrl.dry_dt_ticks = dry_dt_ticks
        ctrl.cfl_safety = self.cfl_safety
        planned = False
        finite = True
        wall0 = _time.monotonic()
        last_log = wall0
        for stop in output_window_stops(start_ticks, end_ticks, interval_ticks):
            ctrl.stop_ticks = stop
            self._ctx.ctrl_write(ctrl)
            if planned:
                self._ctx.retarget()
            else:
                self._ctx.prologue()
                planned = True
            ctrl = self._ctx.ctrl_read()
            while int(ctrl.now_ticks) < stop:
                before = int(ctrl.now_ticks)
                self._ctx.steps(self.steps_per_graph, use_graph=True)
                ctrl = self._ctx.ctrl_read()
                if int(ctrl.now_ticks) == before:
                    raise RuntimeError(
                        f"GPU-resident run made no time progress at "
                        f"t={ticks_to_seconds(before):.6f}s (CFL bound below one tick)"
                    )
                now = _time.monotonic()
                if progress and now - last_log >= 10.0:
                    sys.stdout.write(
                        f"[solver][hip-resident {int(ctrl.steps)}] "
                        f"sim={ticks_to_seconds(int(ctrl.now_ticks)):.2f}/{ticks_to_seconds(end_ticks):.2f}s\n"
                    )
                    sys.stdout.flush()
                    last_log = now
            h_snap = self.download_h()
            if not np.isfinite(h_snap).all():
                snapshots.append(np.nan_to_num(h_snap, nan=0.0, posinf=0.0))
                snap_times.append(ticks_to_seconds(stop))
                capture()
                finite = False
                break
            if (stop - start_ticks) % interval_ticks == 0:
                snapshots.append(h_snap)
                snap_times.append(ticks_to_seconds(stop))
                capture()
        self.steps_total += int(ctrl.steps)
        if finite:
            h_final = self.download_h()
            if not np.allclose(snapshots[-1], h_final):
                snapshots.append(h_final)
                snap_times.append(ticks_to_seconds(int(ctrl.now_ticks)))
                capture()
        if capture_momentum_snapshots:
            return snapshots, snap_times, hu_snaps, hv_snaps
        return snapshots, snap_times

    def close(self) -> None:
        self._ctx.close()


def available() -> bool:
    """True when the library is built and a HIP device can be opened."""
    try:
        ctx = L.Context(0)
    except Exception:
        return False
    ctx.close()
    return True


__all__ = ["STEPS_PER_GRAPH_DEFAULT", "SWESolverHipResident", "available", "device_mesh_arrays"]

