"""Open boundary stretches: which boundary edges a line picks, and how an inflow spreads.

Reference ``inundation/mesh/boundaries.py``: a stretch is a polyline in world
coordinates and picks the boundary edges whose midpoints lie within a
tolerance of it (``:25-58``); an inflow's discharge is spread over a
stretch's edges by Manning conveyance at a common water level (``:61-98``).
"""

from __future__ import annotations

from typing import TYPE_CHECKING

import numpy as np

if TYPE_CHECKING:
    from collections.abc import Sequence

    from numpy.typing import NDArray

    from .geometry import MeshGeometry

# Halvings of the level bracket: 80 take a 1 m bracket below float64 spacing.
LEVEL_BISECTIONS = 80


def boundary_edge_midpoints(geom: MeshGeometry) -> tuple[NDArray[np.int64], NDArray[np.float64]]:
    """Boundary (``edge_cellR == -1``) edge indices and their midpoints relative to ``geom.origin``."""
    edges = np.flatnonzero(np.asarray(geom.edge_cellR) == -1)
    cell = np.asarray(geom.edge_cellL)[edges].astype(np.int64)
    side = np.asarray(geom.edge_sideL)[edges].astype(np.int64)
    v0 = np.asarray(geom.cell_vertex_ptr)[cell].astype(np.int64)
    deg = np.asarray(geom.cell_vertex_ptr)[cell + 1].astype(np.int64) - v0
    verts = np.asarray(geom.cell_vertices, dtype=np.float64)
    return edges, 0.5 * (verts[v0 + side] + verts[v0 + (side + 1) % deg])


def distance_to_line(points: NDArray[np.float64], line: NDArray[np.float64]) -> NDArray[np.float64]:
    """Euclidean distance from each point to the nearest segment of a polyline."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    q = np.asarray(line, dtype=np.float64).reshape(-1, 2)
    out = np.full(p.shape[0], np.inf)
    for a, b in zip(q[:-1], q[1:], strict=True):
        d = b - a
        t = np.clip(((p - a) @ d) / max(float(d @ d), 1e-30), 0.0, 1.0)
        foot = a + t[:, None] * d
        out = np.minimum(out, np.hypot(p[:, 0] - foot[:, 0], p[:, 1] - foot[:, 1]))
    return out


def select_boundary_edges(
    geom: MeshGeometry, line: Sequence[tuple[float, float]], tolerance_m: float
) -> NDArray[np.int64]:
    """Boundary edges whose midpoints lie within ``tolerance_m`` of a world-coordinate polyline."""
    if len(line) < 2:
        raise ValueError(f"a stretch needs at least 2 points, got {len(line)}")
    local = np.asarray(line, dtype=np.float64) - np.asarray(geom.origin, dtype=np.float64)
    edges, mid = boundary_edge_midpoints(geom)
    return edges[distance_to_line(mid, local) <= tolerance_m]


def conveyance_split(
    discharge_m3s: float,
    lengths: NDArray[np.float32],
    beds: NDArray[np.float32],
    manning: NDArray[np.float32],
    energy_slope: float,
) -> NDArray[np.float32]:
    """Discharge per unit width on each edge of a section that carries ``discharge_m3s``.

    The section fills to the level at which its Manning conveyance
    ``sum(len * sqrt(S) / n * max(level - bed, 0)^(5/3))`` carries the
    discharge (bisection on the level), so water enters along the low part of
    the section. Each edge takes its share of that conveyance; the result is
    float32, for the kernels, and the edge with the largest share absorbs the
    cast's rounding so that ``sum(q * len)`` is the discharge.
    """
    if discharge_m3s <= 0.0:
        return np.zeros(np.asarray(lengths).size, dtype=np.float32)
    length = np.asarray(lengths, dtype=np.float64)
    bed = np.asarray(beds, dtype=np.float64)
    rate = length * np.sqrt(energy_slope) / np.asarray(manning, dtype=np.float64)

    def carried(level: float) -> NDArray[np.float64]:
        return rate * np.maximum(level - bed, 0.0) ** (5.0 / 3.0)

    lo = float(bed.min())
    hi = lo + 1.0
    while carried(hi).sum() < discharge_m3s:
        hi = lo + 2.0 * (hi - lo)
    for _ in range(LEVEL_BISECTIONS):
        mid = 0.5 * (lo + hi)
        if carried(mid).sum() < discharge_m3s:
            lo = mid
        else:
            hi = mid
    share = carried(hi)
    q = (discharge_m3s * share / share.sum() / length).astype(np.float32)
    k = int(np.argmax(share))
    q[k] += np.float32((discharge_m3s - float((q.astype(np.float64) * length).sum())) / length[k])
    return q


__all__ = ["boundary_edge_midpoints", "conveyance_split", "distance_to_line", "select_boundary_edges"]
