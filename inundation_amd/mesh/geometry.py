"""Polygon-mesh geometry for the shallow-water solver, and Hilbert cell ordering.

The arrays built here fix the numbering every kernel and oracle depends on: cell
order (Hilbert distance, stable), edge order (stable by the smaller adjacent
cell), the per-cell CSR gather order (the cell's polygon side order) and the
FP32 rounding of areas, normals and bed steps. They must come out bit-identical
to the reference's build, so the NumPy operations follow the same sequence of
floating-point operations as ``inundation/mesh/geometry.py``:

- area/centroid: shoelace sums per row (``geometry.py:130-170,329-373``);
- half-edge matching by the sorted vertex-pair key (``geometry.py:173-202``);
- lengths and normals oriented from the left to the right cell (``:205-238``);
- bed steps in float64 (``:241-254``);
- CSR in side order (``:445-468``);
- Hilbert order p=10 (Skilling's transform, ``:505-552``) and the edge
  locality sort (``:566-705``).

Coordinates are relative to ``origin`` (the lower-left corner of the vertices).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import NDArray

from ..tuning import EDGE_LENGTH_EPSILON, MIN_CELL_AREA, POLYGON_CENTROID_EPSILON


@dataclass
class MeshGeometry:
    """Geometry record the solver consumes (reference ``geometry.py:37-122``).

    Cell arrays have N entries, edge arrays E entries, CSR arrays one entry per
    (cell, edge) slot. ``edge_cellR == -1`` marks a boundary (wall) edge.
    ``cell_edge_idx_signed`` packs ``(edge << 1) | is_right``.
    """

    origin: NDArray[np.float64]
    centroid: NDArray[np.float32]
    area: NDArray[np.float32]
    zb: NDArray[np.float32]
    cell_vertices: NDArray[np.float32]
    cell_vertex_ptr: NDArray[np.int32]
    cell_bbox: NDArray[np.float32]
    edge_len: NDArray[np.float32]
    edge_nx: NDArray[np.float32]
    edge_ny: NDArray[np.float32]
    edge_dzL: NDArray[np.float32]
    edge_dzR: NDArray[np.float32]
    edge_cellL: NDArray[np.int32]
    edge_cellR: NDArray[np.int32]
    edge_sideL: NDArray[np.int32]
    edge_sideR: NDArray[np.int32]
    cell_edge_ptr: NDArray[np.int32]
    cell_edge_idx: NDArray[np.int32]
    cell_nbr_idx: NDArray[np.int32]
    cell_edge_side: NDArray[np.int32]
    cell_edge_idx_signed: NDArray[np.int32]
    N: int
    E: int
    V: int
    max_degree: int = 3
    manning_n: NDArray[np.float32] | None = None

    @property
    def world_centroid(self) -> NDArray[np.float64]:
        """Cell centroids in world coordinates."""
        return self.centroid.astype(np.float64) + self.origin


def _shoelace(xy: NDArray[np.float64]) -> tuple[NDArray[np.float64], NDArray[np.float64]]:
    """Per-row cross products and signed areas of polygons given as (n, d, 2)."""
    x, y = xy[:, :, 0], xy[:, :, 1]
    cross = x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y
    return cross, 0.5 * np.sum(cross, axis=1)


def _centroids(xy: NDArray[np.float64], cross: NDArray[np.float64], six_area: Any) -> Any:
    x, y = xy[:, :, 0], xy[:, :, 1]
    cx = np.sum((x + np.roll(x, -1, axis=1)) * cross, axis=1) / six_area
    cy = np.sum((y + np.roll(y, -1, axis=1)) * cross, axis=1) / six_area
    return cx, cy


def _uniform_degree_cells(
    verts: NDArray[np.float64], faces: NDArray[np.int32], n: int, d: int
) -> tuple[NDArray[np.float32], NDArray[np.float64]]:
    """Area and centroid when every cell has ``d`` sides; reorients CW cells in place."""
    fm = faces.reshape(n, d)
    xy = verts[fm]
    cross, signed = _shoelace(xy)
    cw = signed < 0
    if cw.any():
        fm[cw] = fm[cw, ::-1]
        xy = verts[fm]
        cross, signed = _shoelace(xy)
    area = np.maximum(signed, 1e-30).astype(np.float32)
    cx, cy = _centroids(xy, cross, 6.0 * signed)
    degen = np.abs(signed) < POLYGON_CENTROID_EPSILON
    if degen.any():
        cx[degen] = xy[degen, :, 0].mean(axis=1)
        cy[degen] = xy[degen, :, 1].mean(axis=1)
    return area, np.column_stack([cx, cy])


def _mixed_degree_cells(
    verts: NDArray[np.float64],
    faces: NDArray[np.int32],
    offsets: NDArray[np.int32],
    degrees: NDArray[np.int32],
) -> tuple[NDArray[np.float32], NDArray[np.float64]]:
    """Area and centroid one degree class at a time; reorients CW cells in place.

    A clockwise cell keeps the area of its original vertex order (its magnitude)
    and takes the centroid of the reversed order, as the reference does.
    """
    area = np.empty(degrees.size, dtype=np.float32)
    centroid = np.empty((degrees.size, 2), dtype=np.float64)
    for d in np.unique(degrees).tolist():
        cells = np.flatnonzero(degrees == d)
        slots = offsets[cells, None].astype(np.int64) + np.arange(d)
        cf = faces[slots]
        cross, signed = _shoelace(verts[cf])
        cw = signed < 0
        if cw.any():
            cf[cw] = cf[cw, ::-1]
            faces[slots] = cf
            cross, _ = _shoelace(verts[cf])
            signed = np.abs(signed)
        area[cells] = np.maximum(signed, 1e-30)
        with np.errstate(divide="ignore", invalid="ignore"):
            cx, cy = _centroids(verts[cf], cross, 6.0 * signed)
        centroid[cells, 0] = cx
        centroid[cells, 1] = cy
    return area, centroid


def _pair_sides(
    faces: NDArray[np.int32], offsets: NDArray[np.int32], n_verts: int
) -> tuple[NDArray[np.int64], NDArray[np.int64], NDArray[np.int64]]:
    """Match polygon sides sharing a vertex pair; edges numbered by their first side.

    Returns (first side, second side or -1, position of each side's next vertex).
    """
    off = offsets.astype(np.int64)
    nxt = np.arange(1, faces.size + 1, dtype=np.int64)
    nxt[off[1:] - 1] = off[:-1]
    a = faces.astype(np.int64)
    b = a[nxt]
    key = np.minimum(a, b) * (n_verts + 1) + np.maximum(a, b)
    del a, b
    order = np.argsort(key, kind="stable")
    key = key[order]
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    del key
    first = order[starts]
    two = np.diff(np.r_[starts, order.size]) >= 2
    second = np.full(starts.size, -1, dtype=np.int64)
    second[two] = order[starts[two] + 1]
    by_first = np.argsort(first)
    return first[by_first], second[by_first], nxt


def _lengths_normals(
    pa: NDArray[np.float64],
    pb: NDArray[np.float64],
    centroid: NDArray[np.float64],
    cell_l: NDArray[np.int32],
    cell_r: NDArray[np.int32],
) -> tuple[NDArray[np.float32], NDArray[np.float32], NDArray[np.float32]]:
    """Edge length and unit normal pointing from the left cell to the right one."""
    dx = pb[:, 0] - pa[:, 0]
    dy = pb[:, 1] - pa[:, 1]
    length = np.sqrt(dx * dx + dy * dy)
    with np.errstate(divide="ignore", invalid="ignore"):
        nx = dy / length
        ny = -dx / length
    tiny = length < EDGE_LENGTH_EPSILON
    nx[tiny] = 0.0
    ny[tiny] = 0.0
    inner = np.flatnonzero(cell_r != -1)
    cl = centroid[cell_l[inner]]
    step = centroid[cell_r[inner]] - cl
    inx, iny = nx[inner], ny[inner]
    dot = step[:, 0] * inx + step[:, 1] * iny
    flip = dot < 0
    tie = np.flatnonzero(dot == 0.0)
    if tie.size:
        mid = 0.5 * (pa[inner[tie]] + pb[inner[tie]])
        off = (mid[:, 0] - cl[tie, 0]) * inx[tie] + (mid[:, 1] - cl[tie, 1]) * iny[tie]
        flip[tie] = off < 0
    fl = inner[flip]
    nx[fl] = -nx[fl]
    ny[fl] = -ny[fl]
    return length.astype(np.float32), nx.astype(np.float32), ny.astype(np.float32)


def edge_bed_steps(
    zb: NDArray[np.float32], edge_cellL: NDArray[np.int32], edge_cellR: NDArray[np.int32]
) -> tuple[NDArray[np.float32], NDArray[np.float32]]:
    """Face elevation minus each side's bed, ``(dzL, dzR)``, both ``>= 0``.

    In float64 from the float32 bed, so the step is exact far above the datum;
    a boundary edge mirrors its left cell, so both steps are zero
    (reference ``geometry.py:241-254``).
    """
    z = zb.astype(np.float64)
    zl = z[edge_cellL]
    zr = np.where(edge_cellR >= 0, z[np.maximum(edge_cellR, 0)], zl)
    face = np.maximum(zl, zr)
    return (face - zl).astype(np.float32), (face - zr).astype(np.float32)


def build_geometry(
    verts: NDArray[np.floating[Any]],
    faces: NDArray[np.int32],
    face_offsets: NDArray[np.int32] | None = None,
    *,
    zb_from_file: NDArray[np.float32],
    manning_n_from_file: NDArray[np.float32] | None = None,
    progress: bool = False,
) -> MeshGeometry:
    """Precompute all solver geometry from vertices and polygon faces.

    ``faces`` is either (N, d) for uniform-degree meshes or a flat array with
    ``face_offsets`` (N+1). Same contract and errors as reference
    ``geometry.py:262-502``.
    """
    del progress
    world = np.asarray(verts, dtype=np.float64)
    origin = world.min(axis=0)
    vxy = world - origin
    del world
    faces = np.array(faces, dtype=np.int32, copy=True)
    if faces.ndim == 2:
        n, d = (int(s) for s in faces.shape)
        face_offsets = np.arange(0, (n + 1) * d, d, dtype=np.int32)
        faces = faces.ravel()
    else:
        if face_offsets is None:
            raise ValueError("face_offsets is required when faces is a 1-D array")
        face_offsets = np.asarray(face_offsets, dtype=np.int32)
        n = len(face_offsets) - 1
    n_verts = len(vxy)

    degrees = np.diff(face_offsets)
    max_degree = int(degrees.max())
    if int(degrees.min()) == max_degree:
        area, centroid = _uniform_degree_cells(vxy, faces, n, max_degree)
    else:
        area, centroid = _mixed_degree_cells(vxy, faces, face_offsets, degrees)
    slivers = np.flatnonzero(area <= MIN_CELL_AREA)
    if slivers.size:
        raise ValueError(
            f"{slivers.size} cell(s) have area <= {MIN_CELL_AREA} m^2 "
            f"(first: {slivers[:5].tolist()}); merge or remove degenerate cells"
        )
    vptr = np.asarray(face_offsets, dtype=np.int32).copy()
    cverts = vxy[faces].astype(np.float32, copy=True)
    bbox = np.column_stack(
        [
            np.minimum.reduceat(cverts[:, 0], vptr[:-1]),
            np.maximum.reduceat(cverts[:, 0], vptr[:-1]),
            np.minimum.reduceat(cverts[:, 1], vptr[:-1]),
            np.maximum.reduceat(cverts[:, 1], vptr[:-1]),
        ]
    ).astype(np.float32)

    zb = np.asarray(zb_from_file, dtype=np.float32)
    if zb.shape != (n,):
        raise ValueError(f"zb_from_file has shape {zb.shape}, expected ({n},)")
    if not np.isfinite(zb).all():
        raise ValueError("zb_from_file must contain only finite values")
    manning_n = None
    if manning_n_from_file is not None:
        manning_n = np.asarray(manning_n_from_file, dtype=np.float32).reshape(-1)
        if manning_n.size != n:
            raise ValueError(f"manning_n: expected {n} values, got {manning_n.size}")
        if not np.isfinite(manning_n).all() or not (manning_n > 0.0).all():
            raise ValueError("manning_n must contain finite positive values")

    first, second, nxt = _pair_sides(faces, face_offsets, n_verts)
    n_edges = first.size
    slot_cell = np.repeat(np.arange(n, dtype=np.int32), degrees)
    slot_side = (np.arange(faces.size) - face_offsets[slot_cell].astype(np.int64)).astype(np.int32)
    inner = np.flatnonzero(second >= 0)
    cell_l = slot_cell[first]
    cell_r = np.full(n_edges, -1, dtype=np.int32)
    cell_r[inner] = slot_cell[second[inner]]
    side_l = slot_side[first]
    side_r = np.full(n_edges, -1, dtype=np.int32)
    side_r[inner] = slot_side[second[inner]]
    elen, enx, eny = _lengths_normals(vxy[faces[first]], vxy[faces[nxt[first]]], centroid, cell_l, cell_r)
    del first, second, nxt, slot_cell, slot_side

    m = cell_r != -1
    step = centroid[cell_r[m]] - centroid[cell_l[m]]
    if not ((step[:, 0] * enx[m] + step[:, 1] * eny[m]) >= 0).all():
        raise ValueError("Some interior edge normals point the wrong way!")
    dzl, dzr = edge_bed_steps(zb, cell_l, cell_r)

    slot_cells = np.r_[cell_l, cell_r[inner]]
    if slot_cells.size >= 2**31:
        raise ValueError(f"mesh too large for int32 CSR offsets: total_slots={slot_cells.size}")
    slot_sides = np.r_[side_l, side_r[inner]]
    order = np.argsort(slot_cells.astype(np.int64) * max_degree + slot_sides)
    del slot_sides
    ptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(slot_cells, minlength=n), out=ptr[1:])
    del slot_cells
    cidx = np.r_[np.arange(n_edges, dtype=np.int32), inner.astype(np.int32)][order]
    cnbr = np.r_[cell_r, cell_l[inner]][order]
    cside = np.r_[np.zeros(n_edges, dtype=np.int32), np.ones(inner.size, dtype=np.int32)][order]
    del order
    return MeshGeometry(
        origin=origin,
        centroid=centroid.astype(np.float32),
        area=area,
        zb=zb,
        manning_n=manning_n,
        cell_vertices=cverts,
        cell_vertex_ptr=vptr,
        cell_bbox=bbox,
        edge_len=elen,
        edge_nx=enx,
        edge_ny=eny,
        edge_dzL=dzl,
        edge_dzR=dzr,
        edge_cellL=cell_l,
        edge_cellR=cell_r,
        edge_sideL=side_l,
        edge_sideR=side_r,
        cell_edge_ptr=ptr,
        cell_edge_idx=cidx,
        cell_nbr_idx=cnbr,
        cell_edge_side=cside,
        cell_edge_idx_signed=(cidx.astype(np.int32) << 1) | cside,
        N=n,
        E=n_edges,
        V=n_verts,
        max_degree=max_degree,
    )


def hilbert_distances(
    x: NDArray[np.int64], y: NDArray[np.int64], hilbert_p: int
) -> NDArray[np.int64]:
    """Distance along the order-``hilbert_p`` Hilbert curve of integer points.

    Skilling's axes-to-transpose transform (J. Skilling, "Programming the
    Hilbert curve", AIP Conf. Proc. 707, 2004), then bit interleaving with x
    as the more significant axis; the same curve as the ``hilbertcurve``
    package the reference pins (``geometry.py:505-536``).
    """
    a = [x.astype(np.int64), y.astype(np.int64)]
    top = 1 << (hilbert_p - 1)
    # Inverse undo: for each bit from the top down, invert or exchange low bits.
    q = top
    while q > 1:
        low = q - 1
        a[0] = np.where((a[0] & q) != 0, a[0] ^ low, a[0])
        t = (a[0] ^ a[1]) & low
        hit = (a[1] & q) != 0
        a0_new = np.where(hit, a[0] ^ low, a[0] ^ t)
        a[1] = np.where(hit, a[1], a[1] ^ t)
        a[0] = a0_new
        q >>= 1
    # Gray encode.
    a[1] = a[1] ^ a[0]
    t = np.zeros_like(a[1])
    q = top
    while q > 1:
        t = t ^ np.where((a[1] & q) != 0, q - 1, 0)
        q >>= 1
    a[0] = a[0] ^ t
    a[1] = a[1] ^ t
    dist = np.zeros_like(a[0])
    for bit in range(hilbert_p):
        dist |= ((a[0] >> bit) & 1) << (2 * bit + 1)
        dist |= ((a[1] >> bit) & 1) << (2 * bit)
    return dist


def hilbert_permutation(centroid: NDArray[np.float32], hilbert_p: int = 10) -> NDArray[np.int32]:
    """``perm[new] = old``: cells stable-sorted by Hilbert distance of their centroid."""
    grid = 2**hilbert_p - 1
    cx = centroid[:, 0].astype(np.float64)
    cy = centroid[:, 1].astype(np.float64)
    lo = np.array([cx.min(), cy.min()])
    span = np.array([cx.max() - cx.min(), cy.max() - cy.min()])
    span[span == 0] = 1.0
    ix = ((cx - lo[0]) / span[0] * grid).astype(np.int64)
    iy = ((cy - lo[1]) / span[1] * grid).astype(np.int64)
    return np.argsort(hilbert_distances(ix, iy, hilbert_p), kind="stable").astype(np.int32)


def _rows_in_new_order(
    ptr: NDArray[np.int32], perm: NDArray[np.int32]
) -> tuple[NDArray[np.int32], NDArray[np.int64]]:
    """A CSR whose rows are taken in the order ``perm``: its row pointers and the old slot of each new slot.

    Every old slot is labelled with the new position of its row; a stable
    sort on that label lists the slots row by row in the new order, each
    row's slots in their old order.
    """
    count = np.diff(ptr).astype(np.int64)
    new_pos = np.empty(count.size, dtype=np.int64)
    new_pos[np.asarray(perm, dtype=np.int64)] = np.arange(count.size, dtype=np.int64)
    label = np.repeat(new_pos, count)
    old_slot = np.argsort(label, kind="stable").astype(np.int64)
    new_ptr = np.concatenate([[0], np.cumsum(count[perm])]).astype(np.int32)
    return new_ptr, old_slot


def hilbert_reorder(
    geom: MeshGeometry, hilbert_p: int = 10, verbose: bool = True
) -> tuple[MeshGeometry, NDArray[np.int32]]:
    """Renumber cells along the Hilbert curve and edges by their smaller cell.

    Returns the reordered geometry and ``perm[new] = old``. The per-cell CSR
    keeps each cell's side order; the edge sort is stable by
    ``min(cellL, cellR)`` (``cellL`` for boundary edges), reference
    ``geometry.py:566-705``.
    """
    del verbose
    n = geom.N
    perm = hilbert_permutation(geom.centroid, hilbert_p=hilbert_p)
    inv = np.empty(n, dtype=np.int32)
    inv[perm] = np.arange(n, dtype=np.int32)

    centroid = geom.centroid[perm]
    vptr, vslots = _rows_in_new_order(geom.cell_vertex_ptr, perm)
    cverts = geom.cell_vertices[vslots]
    del vslots

    cl = inv[geom.edge_cellL].astype(np.int32)
    cr = geom.edge_cellR.copy()
    inner = geom.edge_cellR != -1
    cr[inner] = inv[geom.edge_cellR[inner]]

    eptr, eslots = _rows_in_new_order(geom.cell_edge_ptr, perm)
    cidx = geom.cell_edge_idx[eslots]
    cside = geom.cell_edge_side[eslots]
    old_nbr = geom.cell_nbr_idx[eslots]
    cnbr = np.where(old_nbr == -1, -1, inv[old_nbr])
    del eslots, old_nbr

    m = cr != -1
    step = centroid[cr[m]] - centroid[cl[m]]
    dots = step[:, 0] * geom.edge_nx[m] + step[:, 1] * geom.edge_ny[m]
    if not (dots >= -2.0 * np.spacing(centroid.max())).all():
        raise ValueError("hilbert_reorder: an edge normal no longer points from its left cell to its right cell")

    key = np.where(cr >= 0, np.minimum(cl, cr), cl)
    eperm = np.argsort(key, kind="stable").astype(np.int32)
    einv = np.empty(geom.E, dtype=np.int32)
    einv[eperm] = np.arange(geom.E, dtype=np.int32)
    cidx_new = einv[cidx]
    out = MeshGeometry(
        origin=geom.origin,
        centroid=centroid,
        area=geom.area[perm],
        zb=geom.zb[perm],
        manning_n=None if geom.manning_n is None else geom.manning_n[perm],
        cell_vertices=cverts,
        cell_vertex_ptr=vptr,
        cell_bbox=geom.cell_bbox[perm],
        edge_len=geom.edge_len[eperm],
        edge_nx=geom.edge_nx[eperm],
        edge_ny=geom.edge_ny[eperm],
        edge_dzL=geom.edge_dzL[eperm],
        edge_dzR=geom.edge_dzR[eperm],
        edge_cellL=cl[eperm],
        edge_cellR=cr[eperm],
        edge_sideL=geom.edge_sideL[eperm],
        edge_sideR=geom.edge_sideR[eperm],
        cell_edge_ptr=eptr,
        cell_edge_idx=cidx_new,
        cell_nbr_idx=cnbr,
        cell_edge_side=cside,
        cell_edge_idx_signed=(cidx_new.astype(np.int32) << 1) | cside,
        N=n,
        E=geom.E,
        V=geom.V,
        max_degree=geom.max_degree,
    )
    return out, perm


__all__ = [
    "MeshGeometry",
    "build_geometry",
    "edge_bed_steps",
    "hilbert_distances",
    "hilbert_permutation",
    "hilbert_reorder",
]
