"""Mesh geometry, Hilbert ordering, active-set block topology and synthetic meshes."""

from .blocks import EdgeGrouping, edge_order_by_block, group_edges_by_block, num_blocks, remap_cell_edges
from .geometry import (
    MeshGeometry,
    build_geometry,
    edge_bed_steps,
    hilbert_distances,
    hilbert_permutation,
    hilbert_reorder,
)

__all__ = [
    "EdgeGrouping",
    "MeshGeometry",
    "build_geometry",
    "edge_bed_steps",
    "edge_order_by_block",
    "group_edges_by_block",
    "hilbert_distances",
    "hilbert_permutation",
    "hilbert_reorder",
    "num_blocks",
    "remap_cell_edges",
]
