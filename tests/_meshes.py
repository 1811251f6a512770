"""Test meshes and initial states shared by the test modules.

The valley follows the reference's test problem (``tests/unit/_valley.py:19-71``):
quads of DX metres whose bed falls 0.002 in +x with a Gaussian valley along
the centre line, Hilbert-ordered. The four step-parity cases follow
``tests/unit/test_gpu_step_parity.py:81-115``.
"""

from __future__ import annotations

import numpy as np

from inundation_amd.mesh.geometry import MeshGeometry, build_geometry, hilbert_reorder
from inundation_amd.mesh.synth import quad_grid

MANNING = 0.03
DRY_TOL = 1e-3
DT = 2.0**-10


def valley_grid(nx: int = 256, ny: int = 128, dx: float = 2.0, datum: float = 0.0) -> MeshGeometry:
    verts, faces = quad_grid(nx, ny, dx)
    verts = verts.astype(np.float32)
    cx = (np.arange(nx * ny) % nx + 0.5) * dx
    cy = (np.arange(nx * ny) // nx + 0.5) * dx
    zb = (datum - 0.002 * cx - 0.3 * np.exp(-(((cy - ny * dx / 2) / (0.2 * ny * dx)) ** 2))).astype(np.float32)
    geom = build_geometry(verts, faces, zb_from_file=zb)
    return hilbert_reorder(geom, verbose=False)[0]


def channel_grid(
    cells: int = 300, across: int = 20, length: float = 3000.0, width: float = 200.0, slope: float = 1e-3
) -> MeshGeometry:
    """A straight channel of square cells, bed ``10 - slope * x``, Hilbert-ordered.

    The reference's open-boundary test channel (``tests/unit/test_gpu_open_boundaries.py:41-56``:
    3 km by 200 m in 300 by 20 cells, slope 1e-3).
    """
    dx = length / cells
    assert abs(dx - width / across) < 1e-9, "square cells"
    verts, faces = quad_grid(cells, across, dx)
    cx = (np.arange(cells * across) % cells + 0.5) * dx
    geom = build_geometry(verts.astype(np.float32), faces, zb_from_file=(10.0 - slope * cx).astype(np.float32))
    return hilbert_reorder(geom, verbose=False)[0]


def step_shore() -> tuple[MeshGeometry, np.ndarray]:
    """A 32 x 32 basin of unit cells in row order, half a 1 m deep pool at
    rest and half a bank 2 m high: 256-cell blocks 0 and 1 are wet, 2 and 3
    dry, and the last row of block 1 borders the dry bank."""
    verts, faces = quad_grid(32, 32, 1.0)
    cy = verts[faces].mean(axis=1)[:, 1]
    g = build_geometry(verts, faces, zb_from_file=np.where(cy < 16, 0.0, 2.0).astype(np.float32))
    return g, np.where(g.zb == 0.0, 1.0, 0.0).astype(np.float32)


def channel_end(x: float, width: float = 200.0, dx: float = 10.0) -> tuple[list[tuple[float, float]], float]:
    """The line across the channel at ``x`` and its pick tolerance (a quarter cell)."""
    return [(x, 0.0), (x, width)], 0.25 * dx


def _extent(g: MeshGeometry) -> tuple[float, float]:
    v = g.cell_vertices
    return float(v[:, 0].max()), float(v[:, 1].max())


def dam(g: MeshGeometry) -> np.ndarray:
    lx, ly = _extent(g)
    x, y = g.centroid[:, 0], g.centroid[:, 1]
    inside = (x < lx * 0.25) & (np.abs(y - ly / 2) < ly * 0.15)
    return np.where(inside, 1.0, 0.0).astype(np.float32)


def sheet(g: MeshGeometry) -> np.ndarray:
    lx, _ = _extent(g)
    return np.where(g.centroid[:, 0] < lx * 0.3, 0.02, 0.0).astype(np.float32)


def wet_lake(g: MeshGeometry) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    lx, _ = _extent(g)
    x = g.centroid[:, 0].astype(np.float64)
    h = float(g.zb.max()) + 0.5 - g.zb.astype(np.float64)
    hu = 0.1 * np.sin(2 * np.pi * x / lx) * h
    return h, hu, np.zeros_like(h)


def dam_break(g: MeshGeometry) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    h = dam(g).astype(np.float64)
    return h, np.zeros_like(h), np.zeros_like(h)


def draining(g: MeshGeometry) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    h = sheet(g).astype(np.float64)
    return h, np.where(h > 0.0, 0.05, 0.0), np.zeros_like(h)


def dry_momentum(g: MeshGeometry) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    lx, _ = _extent(g)
    h = dam(g).astype(np.float64)
    hu = np.where(g.centroid[:, 0] > lx * 0.8, 0.1, 0.0)
    return h, hu, -hu


CASES = {
    "wet_lake": wet_lake,
    "dam_break": dam_break,
    "draining": draining,
    "dry_momentum": dry_momentum,
}


def f32_state(h, hu, hv):  # noqa: ANN001, ANN201
    return tuple(np.asarray(a, dtype=np.float32) for a in (h, hu, hv))


def split_sides(verts: np.ndarray, faces: np.ndarray, offsets: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Every polygon side cut at its midpoint (one shared vertex per side): twice the edges and CSR slots per cell.

    Two collinear edges then join each pair of neighbours; the mesh stays
    conforming. Dense blocks of such cells exceed what the stage kernels'
    tile levels preload, so the loops after their first barrier run.
    """
    deg = np.diff(offsets)
    cell = np.repeat(np.arange(deg.size), deg)
    pos = np.arange(faces.size) - offsets[cell]
    nxt = faces[offsets[cell] + (pos + 1) % deg[cell]]
    a, b = np.minimum(faces, nxt), np.maximum(faces, nxt)
    key, inv = np.unique(a.astype(np.int64) * (verts.shape[0] + 1) + b, return_inverse=True)
    ka, kb = key // (verts.shape[0] + 1), key % (verts.shape[0] + 1)
    mids = 0.5 * (verts[ka] + verts[kb])
    mid_id = verts.shape[0] + inv
    out = np.empty(2 * faces.size, dtype=np.int32)
    out[0::2] = faces
    out[1::2] = mid_id
    return np.concatenate([verts, mids]), out, (2 * offsets).astype(np.int32)


def dense_voronoi(n_points: int = 12_000, seed: int = 31, length: float = 600.0) -> MeshGeometry:
    """A Voronoi mesh with split sides over fBm terrain, Hilbert-ordered (about 12 sides per cell)."""
    from inundation_amd.mesh.synth import fbm_terrain, voronoi_mesh

    rng = np.random.default_rng(seed)
    verts, faces, offsets = voronoi_mesh(n_points, length, rng)
    verts, faces, offsets = split_sides(np.asarray(verts, dtype=np.float64), faces, offsets)
    deg = np.diff(offsets)
    cell = np.repeat(np.arange(deg.size), deg)
    cx = np.bincount(cell, verts[faces, 0], minlength=deg.size) / deg
    cy = np.bincount(cell, verts[faces, 1], minlength=deg.size) / deg
    zb = fbm_terrain(cx, cy, length, rng)
    return hilbert_reorder(build_geometry(verts, faces, offsets, zb_from_file=zb), verbose=False)[0]
