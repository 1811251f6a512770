"""Geometry cache: a built, reordered ``MeshGeometry`` on disk.

Reference ``inundation/mesh/cache.py:40-292``. Two uses:

- a key-addressed directory cache (``WorkflowConfig.geometry_cache_dir``):
  the key hashes the cache version, the reorder mode and the mesh file's
  bytes; any problem reading an entry is a miss and the geometry is rebuilt;
- explicit artifacts (``build_geometry_cache`` / ``load_geometry_cache_artifact``,
  ``WorkflowConfig.geometry_cache_source``): the geometry plus provenance,
  which stand in for the mesh file; any problem raises.

The file is an ``.npz`` of the geometry's named arrays and scalars, the
permutation ``perm`` (``perm[new] = old``) and, for artifacts, a JSON
``meta`` string; the format version is 6. Field names, the version, the
meta keys and the key hash are the reference's, so caches written by either
load in the other. Writes go to a temporary file that is then renamed.
"""

from __future__ import annotations

import hashlib
import json
import os
import zipfile
from dataclasses import dataclass
from pathlib import Path
from typing import TYPE_CHECKING, Any

import numpy as np

from .geometry import MeshGeometry, build_geometry, hilbert_reorder
from .loader import load_mesh_file

if TYPE_CHECKING:
    from numpy.typing import NDArray

# Bumped whenever the geometry arrays change shape, dtype or meaning.
GEOMETRY_CACHE_VERSION = 6

GEOMETRY_ARRAY_FIELDS = (
    "origin",
    "centroid",
    "area",
    "zb",
    "manning_n",
    "cell_vertices",
    "cell_vertex_ptr",
    "cell_bbox",
    "edge_len",
    "edge_nx",
    "edge_ny",
    "edge_dzL",
    "edge_dzR",
    "edge_cellL",
    "edge_cellR",
    "edge_sideL",
    "edge_sideR",
    "cell_edge_ptr",
    "cell_edge_idx",
    "cell_nbr_idx",
    "cell_edge_side",
    "cell_edge_idx_signed",
)
GEOMETRY_SCALAR_FIELDS = ("N", "E", "V", "max_degree")
_META = "meta"
_READ_ERRORS = (OSError, ValueError, KeyError, EOFError, zipfile.BadZipFile)


@dataclass(frozen=True)
class GeometryCacheMeta:
    """Provenance of an artifact: writer's cache version, mesh file, cell order."""

    version: int
    mesh_source: str
    use_hilbert_reorder: bool


def geometry_cache_key(mesh_path: str | os.PathLike[str], *, use_hilbert_reorder: bool) -> str:
    """BLAKE2b-256 (hex) of ``v<version>|hilbert=<bool>|`` followed by the mesh file's bytes."""
    h = hashlib.blake2b(digest_size=32)
    h.update(f"v{GEOMETRY_CACHE_VERSION}|hilbert={use_hilbert_reorder}|".encode())
    with Path(mesh_path).open("rb") as f:
        while chunk := f.read(1 << 20):
            h.update(chunk)
    return h.hexdigest()


def geometry_cache_path(cache_dir: str | os.PathLike[str], cache_key: str) -> Path:
    """``<cache_dir>/geometry_<key>.npz``."""
    return Path(cache_dir) / f"geometry_{cache_key}.npz"


def geometry_from_mesh_file(
    mesh_source: str | os.PathLike[str], *, use_hilbert_reorder: bool = True, progress: bool = False
) -> tuple[MeshGeometry, NDArray[np.int32]]:
    """Load a mesh file, build its geometry and (optionally) Hilbert-reorder it."""
    verts, faces, offsets, zb, n = load_mesh_file(mesh_source)
    geom = build_geometry(verts, faces, face_offsets=offsets, zb_from_file=zb, manning_n_from_file=n, progress=progress)
    perm = np.arange(geom.N, dtype=np.int32)
    if use_hilbert_reorder:
        geom, perm = hilbert_reorder(geom, verbose=progress)
    return geom, perm


def build_geometry_cache(
    mesh_source: str | os.PathLike[str],
    out_path: str | os.PathLike[str],
    *,
    use_hilbert_reorder: bool = True,
    progress: bool = False,
) -> Path:
    """Preprocess a mesh file into a standalone artifact (no GPU needed); returns its path."""
    src = os.fspath(mesh_source)
    geom, perm = geometry_from_mesh_file(src, use_hilbert_reorder=use_hilbert_reorder, progress=progress)
    meta = GeometryCacheMeta(GEOMETRY_CACHE_VERSION, src, use_hilbert_reorder)
    return _write(Path(out_path), geom, perm, meta)


def _read(path: Path) -> dict[str, Any]:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load_geometry_cache_artifact(
    path: str | os.PathLike[str],
) -> tuple[MeshGeometry, NDArray[np.int32], GeometryCacheMeta]:
    """Geometry, permutation and provenance of an artifact; raises on any problem."""
    p = Path(path)
    if not p.exists():
        raise FileNotFoundError(f"geometry cache artifact not found: {p}")
    try:
        payload = _read(p)
    except _READ_ERRORS as exc:
        raise ValueError(f"unreadable geometry cache artifact {p}: {exc}") from exc
    need = (*GEOMETRY_ARRAY_FIELDS, *GEOMETRY_SCALAR_FIELDS, "perm", _META)
    missing = [k for k in need if k not in payload]
    if missing:
        raise ValueError(f"geometry cache artifact {p} is missing fields: {missing}")
    meta = _meta_from(payload[_META])
    if meta.version != GEOMETRY_CACHE_VERSION:
        raise ValueError(
            f"geometry cache artifact {p} was written with cache version {meta.version}, this build "
            f"expects {GEOMETRY_CACHE_VERSION} — rebuild it with build_geometry_cache()"
        )
    return _geometry(payload), payload["perm"], meta


def save_geometry_cache(
    cache_dir: str | os.PathLike[str], cache_key: str, geom: MeshGeometry, perm: NDArray[np.int32]
) -> Path:
    """Store ``geom`` and ``perm`` under their key in the directory cache."""
    return _write(geometry_cache_path(cache_dir, cache_key), geom, perm)


def load_geometry_cache(
    cache_dir: str | os.PathLike[str], cache_key: str
) -> tuple[MeshGeometry, NDArray[np.int32]] | None:
    """A directory-cache entry, or None on any miss (absent, truncated, incomplete)."""
    p = geometry_cache_path(cache_dir, cache_key)
    if not p.exists():
        return None
    try:
        payload = _read(p)
    except _READ_ERRORS:
        return None
    if any(k not in payload for k in (*GEOMETRY_ARRAY_FIELDS, *GEOMETRY_SCALAR_FIELDS, "perm")):
        return None
    return _geometry(payload), payload["perm"]


def _write(path: Path, geom: MeshGeometry, perm: NDArray[np.int32], meta: GeometryCacheMeta | None = None) -> Path:
    payload: dict[str, Any] = {}
    for k in GEOMETRY_ARRAY_FIELDS:
        v = getattr(geom, k)
        # manning_n is the one optional array: None is stored empty.
        payload[k] = np.empty(0, dtype=np.float32) if v is None else v
    for k in GEOMETRY_SCALAR_FIELDS:
        payload[k] = np.int64(getattr(geom, k))
    payload["perm"] = perm
    if meta is not None:
        payload[_META] = np.asarray(
            json.dumps(
                {"version": meta.version, "mesh_source": meta.mesh_source, "use_hilbert_reorder": meta.use_hilbert_reorder}
            )
        )
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(path.name + ".tmp")
    with tmp.open("wb") as f:
        np.savez(f, **payload)
    tmp.replace(path)
    return path


def _geometry(payload: dict[str, Any]) -> MeshGeometry:
    arrays = {k: payload[k] for k in GEOMETRY_ARRAY_FIELDS}
    if arrays["manning_n"].size == 0:
        arrays["manning_n"] = None
    return MeshGeometry(
        **arrays,
        N=int(payload["N"]),
        E=int(payload["E"]),
        V=int(payload["V"]),
        max_degree=int(payload["max_degree"]),
    )


def _meta_from(raw: NDArray[Any]) -> GeometryCacheMeta:
    try:
        d = json.loads(str(raw.item()))
        return GeometryCacheMeta(int(d["version"]), str(d["mesh_source"]), bool(d["use_hilbert_reorder"]))
    except (AttributeError, TypeError, ValueError, KeyError) as exc:
        raise ValueError(f"invalid geometry cache metadata: {raw!r}") from exc


__all__ = [
    "GEOMETRY_CACHE_VERSION",
    "GeometryCacheMeta",
    "build_geometry_cache",
    "geometry_cache_key",
    "geometry_cache_path",
    "geometry_from_mesh_file",
    "load_geometry_cache",
    "load_geometry_cache_artifact",
    "save_geometry_cache",
]
