"""Mesh files and the geometry cache (reference ``mesh/loader.py``, ``mesh/cache.py``).

- a geometry-cache artifact written by the reference loads here with every
  array equal to this package's build of the same OBJ, and this package's
  artifact holds the same fields and arrays; the cache key equals the
  reference's for the same file (``tests/golden/artifact_v6_patch.npz``,
  ``mesh_obj_cache.npz``);
- the OBJ loader equals the reference's loader on that file;
- WKB: hand-encoded polygons in both byte orders, with Z, M and an EWKB SRID,
  decode to their rings; vertex welding numbers points by first use (against
  a brute-force loop); a parquet mesh round-trips to the same geometry;
- the cache's failure rules (``tests/unit/test_geometry_cache.py``): a
  missing, truncated, incomplete, stale or badly described artifact raises, a
  bad directory entry is a miss; ``load_geometry`` prefers the artifact,
  checks its cell order and uses the directory cache.
"""

from __future__ import annotations

import struct
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from inundation_amd.mesh import cache as C
from inundation_amd.mesh.geometry import build_geometry
from inundation_amd.mesh.loader import (
    load_mesh_file,
    weld_rings,
    wkb_polygon,
    wkb_polygon_exterior,
    write_mesh_parquet,
)
from inundation_amd.mesh.synth import quad_grid
from inundation_amd.workflow import WorkflowConfig, load_geometry

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def golden():  # noqa: ANN201
    return dict(np.load(GOLDEN / "mesh_obj_cache.npz", allow_pickle=False))


@pytest.fixture
def obj_file(tmp_path: Path, golden) -> Path:  # noqa: ANN001
    p = tmp_path / "patch.obj"
    p.write_bytes(golden["obj_text"].tobytes())
    return p


def _same_geometry(a, b) -> None:  # noqa: ANN001
    for k in C.GEOMETRY_ARRAY_FIELDS:
        x, y = getattr(a, k), getattr(b, k)
        if x is None or y is None:
            assert x is None and y is None, k
            continue
        assert x.dtype == y.dtype and np.array_equal(x, y), k
    for k in C.GEOMETRY_SCALAR_FIELDS:
        assert getattr(a, k) == getattr(b, k), k


def test_obj_loader_matches_the_reference(obj_file: Path, golden) -> None:  # noqa: ANN001
    v, f, o, zb, n = load_mesh_file(obj_file)
    assert n is None
    for got, want in ((v, golden["verts"]), (f, golden["faces"]), (o, golden["offsets"]), (zb, golden["zb"])):
        assert got.dtype == want.dtype and np.array_equal(got, want)


def test_reference_artifact_loads_as_this_build(obj_file: Path, tmp_path: Path) -> None:
    geom, perm, meta = C.load_geometry_cache_artifact(GOLDEN / "artifact_v6_patch.npz")
    assert meta.version == 6 and meta.use_hilbert_reorder
    mine, my_perm = C.geometry_from_mesh_file(obj_file, use_hilbert_reorder=True)
    _same_geometry(geom, mine)
    assert np.array_equal(perm, my_perm)
    # This build's artifact has the reference's fields, arrays and meta keys.
    art = C.build_geometry_cache(obj_file, tmp_path / "mine.npz")
    ref = np.load(GOLDEN / "artifact_v6_patch.npz", allow_pickle=False)
    out = np.load(art, allow_pickle=False)
    assert sorted(ref.files) == sorted(out.files)
    for k in ref.files:
        if k != "meta":
            assert ref[k].dtype == out[k].dtype and np.array_equal(ref[k], out[k]), k
    assert sorted(C._meta_from(out["meta"]).__dict__) == sorted(C._meta_from(ref["meta"]).__dict__)


def test_cache_key_is_the_reference_key(obj_file: Path, golden) -> None:  # noqa: ANN001
    assert C.geometry_cache_key(obj_file, use_hilbert_reorder=True) == str(golden["key_hilbert"])
    assert C.geometry_cache_key(obj_file, use_hilbert_reorder=False) == str(golden["key_plain"])


# ---- WKB and welding --------------------------------------------------------


def _wkb(order: str, code: int, pts: list[tuple[float, ...]], srid: int | None = None) -> bytes:
    head = struct.pack(order + "BI", 1 if order == "<" else 0, code)
    if srid is not None:
        head += struct.pack(order + "I", srid)
    body = struct.pack(order + "II", 1, len(pts))
    for p in pts:
        body += struct.pack(order + "d" * len(p), *p)
    return head + body


SQUARE = [(0.0, 0.0), (2.0, 0.0), (2.0, 1.0), (0.0, 1.0), (0.0, 0.0)]


@pytest.mark.parametrize("order", ["<", ">"])
@pytest.mark.parametrize(
    ("code", "extra"),
    [(3, ()), (1003, (5.0,)), (2003, (7.0,)), (3003, (5.0, 7.0)), (0x80000003, (5.0,)), (0xC0000003, (5.0, 7.0))],
)
def test_wkb_polygons_decode_to_their_ring(order: str, code: int, extra: tuple[float, ...]) -> None:
    pts = [p + extra for p in SQUARE]
    ring = wkb_polygon_exterior(_wkb(order, code, pts))
    assert np.array_equal(ring, np.array(SQUARE[:-1]))


def test_ewkb_with_srid_and_our_encoder() -> None:
    blob = _wkb("<", 0x20000003, SQUARE, srid=4326)
    assert np.array_equal(wkb_polygon_exterior(blob), np.array(SQUARE[:-1]))
    assert wkb_polygon(np.array(SQUARE[:-1])) == _wkb("<", 3, SQUARE)
    with pytest.raises(ValueError, match="not a polygon"):
        wkb_polygon_exterior(struct.pack("<BIdd", 1, 1, 0.0, 0.0))


def test_welding_numbers_points_by_first_use() -> None:
    rng = np.random.default_rng(3)
    pool = rng.integers(0, 6, size=(9, 2)).astype(np.float64)
    rings = [pool[rng.choice(9, size=k, replace=False)] for k in (3, 4, 5, 3, 4)]
    verts, faces, offs = weld_rings(rings)
    seen: dict[tuple[float, float], int] = {}
    want = []
    for r in rings:
        for p in map(tuple, r):
            want.append(seen.setdefault(p, len(seen)))
    assert faces.tolist() == want
    assert np.array_equal(verts, np.array(list(seen)))
    assert offs.tolist() == [0, 3, 7, 12, 15, 19]


def test_parquet_mesh_round_trips(tmp_path: Path) -> None:
    verts, faces = quad_grid(6, 4, 2.5)
    verts = verts + np.array([1000.0, 2000.0])
    zb = np.linspace(1.0, 3.0, faces.shape[0]).astype(np.float32)
    n = np.full(faces.shape[0], 0.04, dtype=np.float32)
    p = tmp_path / "m.parquet"
    write_mesh_parquet(p, verts, faces, zb, n)
    v, f, o, z, nn = load_mesh_file(p)
    assert np.array_equal(z, zb) and np.array_equal(nn, n)
    a = build_geometry(v, f, face_offsets=o, zb_from_file=z, manning_n_from_file=nn)
    b = build_geometry(verts, faces, zb_from_file=zb, manning_n_from_file=n)
    _same_geometry(a, b)


def test_loader_rejects_what_the_reference_rejects(tmp_path: Path) -> None:
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(ValueError, match="without required Z"):
        load_mesh_file(bad)
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2\n")
    with pytest.raises(ValueError, match="need at least 3"):
        load_mesh_file(bad)
    with pytest.raises(ValueError, match="Unsupported format"):
        load_mesh_file(tmp_path / "mesh.xyz")


# ---- cache failure rules and the workflow's geometry source -----------------


def test_artifact_failures_raise(obj_file: Path, tmp_path: Path) -> None:
    with pytest.raises(FileNotFoundError):
        C.load_geometry_cache_artifact(tmp_path / "absent.npz")
    art = C.build_geometry_cache(obj_file, tmp_path / "a.npz")
    torn = tmp_path / "torn.npz"
    torn.write_bytes(art.read_bytes()[:200])
    with pytest.raises(ValueError, match="unreadable"):
        C.load_geometry_cache_artifact(torn)
    payload = dict(np.load(art, allow_pickle=False))
    for change, match in (
        ({"meta": np.asarray('{"version": 5, "mesh_source": "x", "use_hilbert_reorder": true}')}, "cache version 5"),
        ({"meta": np.asarray("not json")}, "invalid geometry cache metadata"),
    ):
        q = tmp_path / "edited.npz"
        np.savez(q, **{**payload, **change})
        with pytest.raises(ValueError, match=match):
            C.load_geometry_cache_artifact(q)
    q = tmp_path / "partial.npz"
    np.savez(q, **{k: v for k, v in payload.items() if k != "edge_len"})
    with pytest.raises(ValueError, match="missing fields"):
        C.load_geometry_cache_artifact(q)


def test_missing_roughness_round_trips(obj_file: Path, tmp_path: Path) -> None:
    geom, perm = C.geometry_from_mesh_file(obj_file)
    assert geom.manning_n is None
    p = C.save_geometry_cache(tmp_path, "k", geom, perm)
    back, _ = C.load_geometry_cache(tmp_path, "k")
    assert back.manning_n is None
    with_n = replace(geom, manning_n=np.full(geom.N, 0.05, np.float32))
    C.save_geometry_cache(tmp_path, "k", with_n, perm)
    back, _ = C.load_geometry_cache(tmp_path, "k")
    assert np.array_equal(back.manning_n, with_n.manning_n)
    assert p.exists() and not p.with_name(p.name + ".tmp").exists()


def test_directory_cache_misses_instead_of_raising(obj_file: Path, tmp_path: Path) -> None:
    assert C.load_geometry_cache(tmp_path, "nothing") is None
    geom, perm = C.geometry_from_mesh_file(obj_file)
    p = C.save_geometry_cache(tmp_path, "k", geom, perm)
    p.write_bytes(p.read_bytes()[:100])
    assert C.load_geometry_cache(tmp_path, "k") is None
    k1 = C.geometry_cache_key(obj_file, use_hilbert_reorder=True)
    k2 = C.geometry_cache_key(obj_file, use_hilbert_reorder=False)
    assert k1 != k2


def _cfg(**kw) -> WorkflowConfig:  # noqa: ANN003
    return WorkflowConfig(manning_n=0.05, output_interval_s=1.0, dt_max=0.1, cfl_interval=1, progress=False, **kw)


def test_config_needs_a_mesh_or_an_artifact() -> None:
    with pytest.raises(ValueError, match="mesh_source is required"):
        _cfg(mesh_source=None)


def test_load_geometry_prefers_the_artifact(obj_file: Path, tmp_path: Path, monkeypatch) -> None:  # noqa: ANN001
    art = C.build_geometry_cache(obj_file, tmp_path / "a.npz")
    import inundation_amd.mesh.cache as cache_mod

    def no_mesh(*a, **k):  # noqa: ANN002, ANN003, ANN202
        raise AssertionError("the mesh file must not be read")

    monkeypatch.setattr(cache_mod, "load_mesh_file", no_mesh)
    geom, perm = load_geometry(_cfg(mesh_source=None, geometry_cache_source=art))
    assert geom.N == 8 and perm.size == 8
    with pytest.raises(ValueError, match="use_hilbert_reorder"):
        load_geometry(_cfg(mesh_source=None, geometry_cache_source=art, use_hilbert_reorder=False))


def test_load_geometry_fills_then_uses_the_directory_cache(obj_file: Path, tmp_path: Path, monkeypatch) -> None:  # noqa: ANN001
    import inundation_amd.mesh.cache as cache_mod

    reads = []
    real = cache_mod.load_mesh_file
    monkeypatch.setattr(cache_mod, "load_mesh_file", lambda p: reads.append(p) or real(p))
    cfg = _cfg(mesh_source=str(obj_file), geometry_cache_dir=tmp_path / "cache")
    a, pa = load_geometry(cfg)
    b, pb = load_geometry(cfg)
    assert len(reads) == 1
    _same_geometry(a, b)
    assert np.array_equal(pa, pb)
    key = C.geometry_cache_key(obj_file, use_hilbert_reorder=True)
    assert C.geometry_cache_path(tmp_path / "cache", key).exists()
