"""Host side of the active set: 256-cell blocks and the edge order the device stores.

A block is ``block_cells`` consecutive (Hilbert-ordered) cells. The device
stores the edges grouped by the block of their left cell, keeping their
order inside each group (the reference's resident backend stores them the
same way, ``inundation/vulkan/gpu_resident_batch.py:173-192``), and the CSR
entries renumbered to that order (``group_edges_by_block`` gives both the
order and each edge's new position).

The block graph and its two-ring (the dilation of the active set) are built
on the device from these arrays (``inu_set_mesh``), so nothing else of the
block topology is needed on the host.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
from numpy.typing import NDArray


def num_blocks(n_cells: int, block_cells: int) -> int:
    """Blocks covering ``n_cells`` cells; the last may be partial."""
    if block_cells < 1:
        raise ValueError(f"block_cells must be positive, got {block_cells}")
    return (n_cells + block_cells - 1) // block_cells


class EdgeGrouping(NamedTuple):
    """``order[new] = old`` and its inverse ``position[old] = new``."""

    order: NDArray[np.int64]
    position: NDArray[np.int64]


def group_edges_by_block(edge_cell_l: NDArray[np.int32], block_cells: int) -> EdgeGrouping:
    """Edges grouped by the block of their left cell, in their own order within a group."""
    blk = np.asarray(edge_cell_l, dtype=np.int64) // block_cells
    order = np.argsort(blk, kind="stable").astype(np.int64)
    position = np.empty_like(order)
    position[order] = np.arange(order.size, dtype=np.int64)
    return EdgeGrouping(order, position)


def edge_order_by_block(edge_cell_l: NDArray[np.int32], block_cells: int) -> NDArray[np.int64]:
    """``order[new] = old``: edges grouped by their left cell's block, stable within a group."""
    return group_edges_by_block(edge_cell_l, block_cells).order


def remap_cell_edges(cell_edge_idx_signed: NDArray[np.int32], order: NDArray[np.int64]) -> NDArray[np.uint32]:
    """CSR entries ``(edge << 1) | is_right`` with edges renumbered by ``order[new] = old``.

    Each cell keeps the order of its entries; only the edge numbers change.
    """
    order = np.asarray(order, dtype=np.int64)
    position = np.empty_like(order)
    position[order] = np.arange(order.size, dtype=np.int64)
    entry = np.asarray(cell_edge_idx_signed, dtype=np.int64)
    return (position[entry >> 1] * 2 + (entry & 1)).astype(np.uint32)


__all__ = ["EdgeGrouping", "edge_order_by_block", "group_edges_by_block", "num_blocks", "remap_cell_edges"]
