"""Polygon mesh files: OBJ and (Geo)Parquet in, (Geo)Parquet out.

Reference ``inundation/mesh/loader.py:81-216`` and ``bench/common.py:891-935``:

- ``.obj``: every ``v`` record carries x, y, z; a cell's bed is the mean of
  its vertex elevations; faces may be any polygon of 3 or more vertices;
- ``.parquet`` / ``.geoparquet``: a WKB ``geometry`` column of simple
  polygons (interior rings ignored, null rows skipped), a required
  ``z_mean`` bed column and an optional ``manning_n`` column. The reference
  decodes WKB with shapely; this package decodes and encodes it itself
  (OGC simple-features WKB and PostGIS EWKB: either byte order, 2D, Z, M or
  ZM points, an optional SRID), so only pyarrow is needed;
- ``.gpkg`` / ``.shp`` go through geopandas when it is installed.

Vertices are welded on their exact float64 coordinates and numbered in the
order they are first met, as the reference's loader numbers them, so the
geometry built from a file is the reference's, bit for bit.
"""

from __future__ import annotations

import os
import struct
from typing import TYPE_CHECKING, Any

import numpy as np

from ..tuning import MIN_POLYGON_VERTICES

if TYPE_CHECKING:
    from pathlib import Path

    from numpy.typing import NDArray

_WKB_POLYGON = 3
_EWKB_Z, _EWKB_M, _EWKB_SRID = 0x80000000, 0x40000000, 0x20000000


def _positive_manning(values: object) -> NDArray[np.float32]:
    try:
        n = np.asarray(values, dtype=np.float32).reshape(-1)
    except (TypeError, ValueError) as exc:
        raise ValueError("manning_n must contain finite positive values") from exc
    if not np.isfinite(n).all() or not (n > 0.0).all():
        raise ValueError("manning_n must contain finite positive values")
    return n


def wkb_polygon_exterior(blob: bytes) -> NDArray[np.float64]:
    """(k, 2) exterior ring of one WKB/EWKB polygon, the closing point dropped."""
    if len(blob) < 9:
        raise ValueError("truncated WKB geometry")
    order = "<" if blob[0] == 1 else ">"
    if blob[0] not in (0, 1):
        raise ValueError("bad WKB byte-order mark")
    (code,) = struct.unpack_from(order + "I", blob, 1)
    pos = 5
    has_z = bool(code & _EWKB_Z)
    has_m = bool(code & _EWKB_M)
    if code & _EWKB_SRID:
        pos += 4
    code &= 0x0FFFFFFF
    # ISO WKB: 1000s add Z, 2000s add M, 3000s add both.
    iso, base = divmod(code, 1000)
    has_z = has_z or iso in (1, 3)
    has_m = has_m or iso in (2, 3)
    if base != _WKB_POLYGON:
        raise ValueError(f"geometry type {code} is not a polygon")
    dims = 2 + int(has_z) + int(has_m)
    (n_rings,) = struct.unpack_from(order + "I", blob, pos)
    pos += 4
    if n_rings == 0:
        return np.empty((0, 2), dtype=np.float64)
    (n_pts,) = struct.unpack_from(order + "I", blob, pos)
    pos += 4
    if pos + 8 * dims * n_pts > len(blob):
        raise ValueError("truncated WKB geometry")
    pts = np.frombuffer(blob, dtype=np.dtype(np.float64).newbyteorder(order), count=dims * n_pts, offset=pos)
    xy = pts.reshape(n_pts, dims)[:, :2].astype(np.float64)
    return xy[:-1] if n_pts > 0 else xy


def wkb_polygon(ring: NDArray[np.float64]) -> bytes:
    """Little-endian 2D WKB polygon of one exterior ring (closed on output)."""
    xy = np.asarray(ring, dtype="<f8").reshape(-1, 2)
    closed = np.vstack([xy, xy[:1]])
    return struct.pack("<BIII", 1, _WKB_POLYGON, 1, closed.shape[0]) + closed.tobytes()


def weld_rings(
    rings: list[NDArray[np.float64]],
) -> tuple[NDArray[np.float64], NDArray[np.int32], NDArray[np.int32]]:
    """Vertices, flat faces and face offsets of polygon rings, welding equal points.

    A vertex's number is the rank of its first use in ring order.
    """
    degrees = np.array([r.shape[0] for r in rings], dtype=np.int64)
    short = np.flatnonzero(degrees < MIN_POLYGON_VERTICES)
    if short.size:
        raise ValueError(f"Geometry has {int(degrees[short[0]])} vertices — need at least 3.")
    if int(degrees.sum()) >= 2**31:
        raise ValueError(f"mesh too large for int32 face offsets: {int(degrees.sum())} polygon vertices")
    offsets = np.zeros(len(rings) + 1, dtype=np.int32)
    np.cumsum(degrees, out=offsets[1:])
    xy = np.concatenate(rings) if rings else np.empty((0, 2), dtype=np.float64)
    uniq, first, inverse = np.unique(xy, axis=0, return_index=True, return_inverse=True)
    rank = np.empty(first.size, dtype=np.int32)
    rank[np.argsort(first, kind="stable")] = np.arange(first.size, dtype=np.int32)
    faces = rank[inverse.ravel()].astype(np.int32)
    verts = np.empty_like(uniq)
    verts[rank] = uniq
    return verts, faces, offsets


def _read_parquet(path: str) -> tuple[Any, ...]:
    import pyarrow.parquet as pq

    table = pq.read_table(path)
    if "geometry" not in table.column_names:
        raise ValueError(f"Mesh file '{path}' has no 'geometry' column")
    blobs = table.column("geometry").to_pylist()
    keep = np.array([b is not None for b in blobs], dtype=bool)
    if not keep.any():
        raise ValueError(
            f"Mesh file '{path}' contains no valid polygon geometries "
            f"({table.num_rows} rows, {int(keep.sum())} non-null geometries)."
        )
    rings = []
    for row in np.flatnonzero(keep):
        try:
            rings.append(wkb_polygon_exterior(blobs[row]))
        except ValueError as exc:
            raise ValueError(f"Mesh file '{path}' row {int(row)}: {exc}") from exc
    verts, faces, offsets = weld_rings(rings)
    if "z_mean" not in table.column_names:
        raise ValueError(f"Mesh file '{path}' is missing required 'z_mean' bed elevations")
    zb = np.asarray(table.column("z_mean").to_numpy(zero_copy_only=False), dtype=np.float64)[keep].astype(np.float32)
    if not np.isfinite(zb).all():
        raise ValueError(f"Mesh file '{path}' contains non-finite 'z_mean' elevations")
    n = None
    if "manning_n" in table.column_names:
        n = _positive_manning(np.asarray(table.column("manning_n").to_numpy(zero_copy_only=False))[keep])
    return verts, faces, offsets, zb, n


def _read_gis(path: str) -> tuple[Any, ...]:
    try:
        import geopandas as gpd
    except ModuleNotFoundError as exc:
        raise RuntimeError(f"reading '{path}' needs geopandas, which is not installed") from exc
    gdf = gpd.read_file(path)
    keep = gdf.geometry.notna().to_numpy()
    if not keep.any():
        raise ValueError(
            f"Mesh file '{path}' contains no valid polygon geometries ({len(gdf)} rows, {int(keep.sum())} non-null geometries)."
        )
    rings = [wkb_polygon_exterior(bytes(g.wkb)) for g in gdf.geometry[keep]]
    verts, faces, offsets = weld_rings(rings)
    if "z_mean" not in gdf.columns:
        raise ValueError(f"Mesh file '{path}' is missing required 'z_mean' bed elevations")
    zb = gdf.loc[keep, "z_mean"].to_numpy(dtype=np.float32)
    if not np.isfinite(zb).all():
        raise ValueError(f"Mesh file '{path}' contains non-finite 'z_mean' elevations")
    n = _positive_manning(gdf.loc[keep, "manning_n"].to_numpy()) if "manning_n" in gdf.columns else None
    return verts, faces, offsets, zb, n


def _read_obj(path: str) -> tuple[Any, ...]:
    xy: list[tuple[float, float]] = []
    z: list[float] = []
    faces: list[list[int]] = []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "v":
                if len(tok) < 4:
                    raise ValueError(f"OBJ file '{path}' has a vertex without required Z elevation")
                xy.append((float(tok[1]), float(tok[2])))
                z.append(float(tok[3]))
            elif tok[0] == "f":
                face = [int(t.split("/")[0]) - 1 for t in tok[1:]]
                if len(face) < MIN_POLYGON_VERTICES:
                    raise ValueError(f"OBJ face has {len(face)} vertices — need at least 3.")
                faces.append(face)
    if not faces:
        raise ValueError(f"OBJ file '{path}' contains no face ('f') records.")
    vz = np.asarray(z, dtype=np.float32)
    if not np.isfinite(vz).all():
        raise ValueError(f"OBJ file '{path}' contains non-finite vertex elevations")
    offsets = np.zeros(len(faces) + 1, dtype=np.int32)
    np.cumsum([len(f) for f in faces], out=offsets[1:])
    flat = np.fromiter((v for f in faces for v in f), dtype=np.int32, count=int(offsets[-1]))
    # Per-face mean of float32 elevations, accumulated in float32 as NumPy's mean does.
    zb = np.asarray([vz[f].mean() for f in faces], dtype=np.float32)
    return np.asarray(xy, dtype=np.float64), flat, offsets, zb, None


def load_mesh_file(
    path: str | os.PathLike[str],
) -> tuple[NDArray[np.float64], NDArray[np.int32], NDArray[np.int32], NDArray[np.float32], NDArray[np.float32] | None]:
    """``(verts, faces, face_offsets, zb, manning_n or None)`` of a polygon mesh file.

    ``verts`` (V, 2) float64 as in the file; face i is
    ``faces[face_offsets[i]:face_offsets[i + 1]]``; ``zb`` per cell.
    """
    p = os.fspath(path)
    ext = os.path.splitext(p)[1].lower()
    if ext in (".parquet", ".geoparquet"):
        return _read_parquet(p)
    if ext in (".gpkg", ".shp"):
        return _read_gis(p)
    if ext == ".obj":
        return _read_obj(p)
    raise ValueError(f"Unsupported format: '{ext}'. Supported: .gpkg, .shp, .parquet, .geoparquet, .obj")


def write_mesh_parquet(
    out_path: str | os.PathLike[str] | Path,
    verts: NDArray[np.floating[Any]],
    faces: NDArray[np.int32],
    zb: NDArray[np.float32],
    manning_n: NDArray[np.float32] | None = None,
) -> None:
    """A polygon mesh as parquet: WKB ``geometry``, ``z_mean`` (float64), optional ``manning_n``."""
    import pyarrow as pa
    import pyarrow.parquet as pq

    faces = np.asarray(faces)
    if faces.ndim != 2:
        raise ValueError(f"faces must be (N, degree), got shape {faces.shape}")
    if np.asarray(zb).shape[0] != faces.shape[0]:
        raise ValueError(f"zb has {np.asarray(zb).shape[0]} rows, expected {faces.shape[0]}")
    v = np.asarray(verts, dtype=np.float64)
    cols = {
        "geometry": pa.array([wkb_polygon(v[f]) for f in faces], type=pa.binary()),
        "z_mean": pa.array(np.asarray(zb).astype(np.float64), type=pa.float64()),
    }
    if manning_n is not None:
        cols["manning_n"] = pa.array(np.asarray(manning_n, dtype=np.float64), type=pa.float64())
    pq.write_table(pa.table(cols), os.fspath(out_path))


__all__ = ["load_mesh_file", "weld_rings", "wkb_polygon", "wkb_polygon_exterior", "write_mesh_parquet"]
