"""Open boundary stretches: which boundary edges a line picks, and how an inflow spreads over them.

The interface follows the reference's ``inundation/mesh/boundaries.py``: a
stretch is a polyline in world coordinates that picks the boundary edges
whose midpoints lie within a tolerance of it, and an inflow's discharge is
spread over the picked edges in proportion to their Manning conveyance at one
common water level. The implementation here is its own:

- the picking computes every midpoint-to-segment distance of the polyline
  in one array (edges x segments) and keeps an edge when the nearest segment
  is within the tolerance;
- the level is found exactly on the piecewise structure of the conveyance:
  the sorted bed elevations bracket it between two consecutive beds (where
  the set of wetted edges is fixed), and Newton's method on that interval,
  started from its upper end where the function is convex, converges
  monotonically to it.
"""

from __future__ import annotations

from typing import TYPE_CHECKING

import numpy as np

if TYPE_CHECKING:
    from collections.abc import Sequence

    from numpy.typing import NDArray

    from .geometry import MeshGeometry

# Manning's conveyance grows as depth^(5/3).
_CONVEYANCE_POWER = 5.0 / 3.0
_NEWTON_STEPS = 60


def boundary_edge_midpoints(geom: MeshGeometry) -> tuple[NDArray[np.int64], NDArray[np.float64]]:
    """Indices of the boundary edges (no right cell) and their midpoints, relative to ``geom.origin``.

    A boundary edge is side ``edge_sideL`` of its left cell's polygon: it runs
    from that polygon's vertex ``side`` to the next one (cyclically).
    """
    cell_r = np.asarray(geom.edge_cellR)
    edges = np.nonzero(cell_r == -1)[0].astype(np.int64)
    ptr = np.asarray(geom.cell_vertex_ptr, dtype=np.int64)
    cells = np.asarray(geom.edge_cellL, dtype=np.int64)[edges]
    start = ptr[cells]
    n_sides = ptr[cells + 1] - start
    side = np.asarray(geom.edge_sideL, dtype=np.int64)[edges]
    corners = np.asarray(geom.cell_vertices, dtype=np.float64)
    tail = corners[start + side]
    head = corners[start + np.where(side + 1 == n_sides, 0, side + 1)]
    return edges, (tail + head) / 2.0


def distance_to_line(points: NDArray[np.float64], line: NDArray[np.float64]) -> NDArray[np.float64]:
    """Distance from each point to the nearest point of a polyline (all segments at once)."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 1, 2)
    q = np.asarray(line, dtype=np.float64).reshape(-1, 2)
    a, b = q[:-1][None, :, :], q[1:][None, :, :]  # (1, segments, 2)
    ab = b - a
    span = np.maximum(np.einsum("psk,psk->ps", ab, ab), 1e-30)
    along = np.clip(np.einsum("psk,psk->ps", p - a, ab) / span, 0.0, 1.0)
    gap = p - (a + along[..., None] * ab)
    return np.sqrt(np.einsum("psk,psk->ps", gap, gap)).min(axis=1)


def select_boundary_edges(
    geom: MeshGeometry, line: Sequence[tuple[float, float]], tolerance_m: float
) -> NDArray[np.int64]:
    """Boundary edges whose midpoints lie within ``tolerance_m`` of a world-coordinate polyline."""
    if len(line) < 2:
        raise ValueError(f"a stretch needs at least 2 points, got {len(line)}")
    edges, mid = boundary_edge_midpoints(geom)
    local = np.asarray(line, dtype=np.float64) - np.asarray(geom.origin, dtype=np.float64)
    near = distance_to_line(mid, local) <= tolerance_m
    return edges[near]


def _conveyance_level(rate: NDArray[np.float64], bed: NDArray[np.float64], discharge: float) -> float:
    """The level L with sum(rate * max(L - bed, 0)^(5/3)) = discharge (discharge > 0)."""
    order = np.argsort(bed, kind="stable")
    b, r = bed[order], rate[order]

    def carried(level: float, k: int) -> tuple[float, float]:
        d = level - b[:k]
        return float((r[:k] * d**_CONVEYANCE_POWER).sum()), float((_CONVEYANCE_POWER * r[:k] * d ** (2.0 / 3.0)).sum())

    # Smallest k such that the level lies in [b[k-1], b[k]): the first bed
    # at which the k lower edges already carry the discharge.
    k = b.size
    for j in range(1, b.size):
        if carried(float(b[j]), j)[0] >= discharge:
            k = j
            break
    lo = float(b[k - 1])
    # Upper start: on [b[k-1], inf) with the k lowest edges wet, the sum is
    # increasing and convex, so Newton from above decreases to the root.
    level = float(b[k]) if k < b.size else lo + 1.0
    while carried(level, k)[0] < discharge:
        level = lo + 2.0 * (level - lo)
    for _ in range(_NEWTON_STEPS):
        f, df = carried(level, k)
        step = (f - discharge) / df if df > 0.0 else 0.0
        nxt = max(level - step, lo)
        if nxt >= level:
            break
        level = nxt
    return level


def conveyance_split(
    discharge_m3s: float,
    lengths: NDArray[np.float32],
    beds: NDArray[np.float32],
    manning: NDArray[np.float32],
    energy_slope: float,
) -> NDArray[np.float32]:
    """Discharge per unit width on each edge of a section that carries ``discharge_m3s``.

    Each edge carries ``len * sqrt(S) / n * max(L - bed, 0)^(5/3)`` at the
    level L where the section carries the whole discharge, so water enters
    along the low part of the section. The result is float32 for the
    kernels; the longest wetted edge takes the float32 rounding, so that
    ``sum(q * len)`` is the discharge.
    """
    length = np.asarray(lengths, dtype=np.float64)
    out = np.zeros(length.size, dtype=np.float32)
    if discharge_m3s <= 0.0 or length.size == 0:
        return out
    bed = np.asarray(beds, dtype=np.float64)
    rate = length * np.sqrt(energy_slope) / np.asarray(manning, dtype=np.float64)
    level = _conveyance_level(rate, bed, float(discharge_m3s))
    share = rate * np.maximum(level - bed, 0.0) ** _CONVEYANCE_POWER
    out[:] = (float(discharge_m3s) / share.sum()) * share / length
    wet = np.nonzero(share > 0.0)[0]
    k = int(wet[np.argmax(length[wet])])
    out[k] += np.float32((float(discharge_m3s) - float((out.astype(np.float64) * length).sum())) / length[k])
    return out


__all__ = ["boundary_edge_midpoints", "conveyance_split", "distance_to_line", "select_boundary_edges"]
