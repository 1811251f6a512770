"""Geometry build, Hilbert order and block helpers are bit-identical to the reference.

The fixtures in ``tests/golden/geometry_*.npz`` hold the output of the
reference's own ``build_geometry`` / ``hilbert_reorder`` (``inundation/mesh/
geometry.py:262-705``) and block helpers (``inundation/mesh/blocks.py``) on
small meshes: uniform quads and triangles, a valley 300 m above the datum,
mixed triangles/quads with every third cell given clockwise, and a hexagon
ring (degrees 3 and 6). Every array must match in dtype and bits.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from inundation_amd.mesh import blocks
from inundation_amd.mesh.geometry import build_geometry, edge_bed_steps, hilbert_reorder
from oracle import blocks_ref

GOLD = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem[len("geometry_") :] for p in GOLD.glob("geometry_*.npz"))
FIELDS = (
    "origin", "centroid", "area", "zb", "cell_vertices", "cell_vertex_ptr", "cell_bbox", "edge_len",
    "edge_nx", "edge_ny", "edge_dzL", "edge_dzR", "edge_cellL", "edge_cellR", "edge_sideL",
    "edge_sideR", "cell_edge_ptr", "cell_edge_idx", "cell_nbr_idx", "cell_edge_side", "cell_edge_idx_signed",
)


def _build(d):  # noqa: ANN001, ANN202
    return build_geometry(
        d["in_verts"],
        d["in_faces"],
        d["in_offsets"] if "in_offsets" in d.files else None,
        zb_from_file=d["in_zb"],
        manning_n_from_file=d["in_manning"] if "in_manning" in d.files else None,
    )


def _same(a: np.ndarray, b: np.ndarray) -> bool:
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return bool(np.array_equal(a.view(f"u{a.dtype.itemsize}"), b.view(f"u{b.dtype.itemsize}")))
    return bool(np.array_equal(a, b))


def test_fixtures_present() -> None:
    assert set(NAMES) >= {"quads2x2", "tri4x4", "valley32x16", "mixed9", "hexring"}


@pytest.mark.parametrize("name", NAMES)
def test_build_geometry_is_bit_identical(name: str) -> None:
    d = np.load(GOLD / f"geometry_{name}.npz")
    g = _build(d)
    for f in FIELDS:
        assert _same(getattr(g, f), d[f"built_{f}"]), f"built {f}"
    assert [g.N, g.E, g.V, g.max_degree] == d["built_counts"].tolist()


@pytest.mark.parametrize("name", NAMES)
def test_hilbert_reorder_is_bit_identical(name: str) -> None:
    d = np.load(GOLD / f"geometry_{name}.npz")
    r, perm = hilbert_reorder(_build(d), verbose=False)
    assert _same(perm, d["perm"])
    for f in FIELDS:
        assert _same(getattr(r, f), d[f"reordered_{f}"]), f"reordered {f}"


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("bc", [2, 4])
def test_block_helpers_are_identical(name: str, bc: int) -> None:
    d = np.load(GOLD / f"geometry_{name}.npz")
    r, _ = hilbert_reorder(_build(d), verbose=False)
    order = blocks.edge_order_by_block(r.edge_cellL, bc)
    assert _same(order, d[f"b{bc}_order"])
    assert _same(blocks.remap_cell_edges(r.cell_edge_idx_signed, order), d[f"b{bc}_remap"])
    assert _same(blocks_ref.edge_workgroup_block_ranges(r.edge_cellL[order], 4, bc), d[f"b{bc}_ranges"])
    ptr, idx = blocks_ref.block_adjacency(r.edge_cellL, r.edge_cellR, r.N, bc)
    assert _same(ptr, d[f"b{bc}_adj_ptr"]) and _same(idx, d[f"b{bc}_adj_idx"])
    p2, i2 = blocks_ref.block_two_ring(ptr, idx)
    assert _same(p2, d[f"b{bc}_ring_ptr"]) and _same(i2, d[f"b{bc}_ring_idx"])
    # One dilation over the two-ring is two over the adjacency, from every single block.
    nb = ptr.size - 1
    for b in range(nb):
        f = np.zeros(nb, dtype=bool)
        f[b] = True
        twice = blocks_ref.dilate_blocks(blocks_ref.dilate_blocks(f, ptr, idx), ptr, idx)
        assert np.array_equal(blocks_ref.dilate_blocks(f, p2, i2), twice)


def test_bed_steps_far_from_the_datum() -> None:
    """Known answer of reference test_mesh_edge_dz.py:69-75."""
    zb = np.array([1000.0, 1000.0 + 2.0**-13, 999.5], dtype=np.float32)
    dzl, dzr = edge_bed_steps(zb, np.array([0, 1, 2], np.int32), np.array([1, 2, -1], np.int32))
    assert np.array_equal(dzl, np.array([2.0**-13, 0.0, 0.0], np.float32))
    assert np.array_equal(dzr, np.array([0.0, 0.5 + 2.0**-13, 0.0], np.float32))


@pytest.mark.parametrize("name", NAMES)
def test_one_side_of_every_face_sits_on_it(name: str) -> None:
    """The face is the higher bed: one step is exactly zero, boundaries both zero."""
    d = np.load(GOLD / f"geometry_{name}.npz")
    r, _ = hilbert_reorder(_build(d), verbose=False)
    assert (np.minimum(r.edge_dzL, r.edge_dzR) == 0.0).all()
    wall = r.edge_cellR < 0
    assert (r.edge_dzL[wall] == 0.0).all() and (r.edge_dzR[wall] == 0.0).all()


@pytest.mark.parametrize("name", NAMES)
def test_signed_csr_is_consistent(name: str) -> None:
    """Each edge has one left slot, interior edges one right slot (test_inundation_signed_csr.py:38-78)."""
    d = np.load(GOLD / f"geometry_{name}.npz")
    r, _ = hilbert_reorder(_build(d), verbose=False)
    cell = np.repeat(np.arange(r.N), np.diff(r.cell_edge_ptr))
    e = r.cell_edge_idx_signed >> 1
    right = (r.cell_edge_idx_signed & 1).astype(bool)
    assert np.array_equal(np.where(right, r.edge_cellR[e], r.edge_cellL[e]), cell)
    assert np.bincount(e[~right], minlength=r.E).tolist() == [1] * r.E
    assert np.array_equal(np.bincount(e[right], minlength=r.E), (r.edge_cellR >= 0).astype(np.int64))


def test_sliver_cells_are_refused() -> None:
    verts = np.array([[0, 0], [1, 0], [2, 0], [1, 1]], dtype=np.float64)
    faces = np.array([[0, 1, 3], [0, 2, 1]], dtype=np.int32)  # second cell is degenerate
    with pytest.raises(ValueError, match="area"):
        build_geometry(verts, faces, zb_from_file=np.zeros(2, np.float32))
