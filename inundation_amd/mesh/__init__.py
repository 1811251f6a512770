"""Mesh geometry, Hilbert ordering, active-set block topology and synthetic meshes."""

from .blocks import (
    block_adjacency,
    block_two_ring,
    dilate_blocks,
    edge_order_by_block,
    edge_workgroup_block_ranges,
    num_blocks,
    remap_cell_edges,
)
from .geometry import (
    MeshGeometry,
    build_geometry,
    edge_bed_steps,
    hilbert_distances,
    hilbert_permutation,
    hilbert_reorder,
)

__all__ = [
    "MeshGeometry",
    "block_adjacency",
    "block_two_ring",
    "build_geometry",
    "dilate_blocks",
    "edge_bed_steps",
    "edge_order_by_block",
    "edge_workgroup_block_ranges",
    "hilbert_distances",
    "hilbert_permutation",
    "hilbert_reorder",
    "num_blocks",
    "remap_cell_edges",
]
