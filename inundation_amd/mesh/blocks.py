"""Host side of the active set: 256-cell blocks and their adjacency.

A block is ``block_cells`` consecutive (Hilbert-ordered) cells. The device
skips every block farther than two block rings from a wet block, so these
helpers build the block graph, its two-ring, the stable edge order by the
left cell's block, and the CSR slot remap that follows that edge order.
``dilate_blocks`` is the host statement of one dilation.

Reference: ``inundation/mesh/blocks.py:21-128``.
"""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray


def num_blocks(n_cells: int, block_cells: int) -> int:
    """Blocks covering ``n_cells`` cells; the last may be partial."""
    if block_cells < 1:
        raise ValueError(f"block_cells must be positive, got {block_cells}")
    return -(-n_cells // block_cells)


def edge_order_by_block(edge_cell_l: NDArray[np.int32], block_cells: int) -> NDArray[np.int64]:
    """Stable permutation ``order[new] = old`` of edges by the block of their left cell."""
    return np.argsort(np.asarray(edge_cell_l) // block_cells, kind="stable")


def remap_cell_edges(
    cell_edge_idx_signed: NDArray[np.int32], order: NDArray[np.int64]
) -> NDArray[np.uint32]:
    """CSR entries ``(edge << 1) | is_right`` renumbered by ``order[new] = old``.

    Each cell keeps the order of its entries; only edge ids change.
    """
    where = np.empty(order.size, dtype=np.int64)
    where[order] = np.arange(order.size, dtype=np.int64)
    s = np.asarray(cell_edge_idx_signed, dtype=np.int64)
    return ((where[s >> 1] << 1) | (s & 1)).astype(np.uint32)


def edge_workgroup_block_ranges(
    sorted_cell_l: NDArray[np.int32], workgroup: int, block_cells: int
) -> NDArray[np.uint32]:
    """``(n_workgroups, 2)`` first and last left-cell block of each run of ``workgroup`` edges."""
    blk = np.asarray(sorted_cell_l, dtype=np.int64) // block_cells
    if blk.size and np.any(np.diff(blk) < 0):
        raise ValueError("edges must be sorted by the block of their left cell")
    lo = np.arange(0, blk.size, workgroup)
    hi = np.minimum(lo + workgroup, blk.size) - 1
    return np.column_stack([blk[lo], blk[hi]]).astype(np.uint32)


def _csr_from_pairs(pairs: NDArray[np.int64], n_rows: int) -> tuple[NDArray[np.uint32], NDArray[np.uint32]]:
    if pairs.size == 0:
        pairs = np.zeros((0, 2), dtype=np.int64)
    ptr = np.zeros(n_rows + 1, dtype=np.uint32)
    np.cumsum(np.bincount(pairs[:, 0], minlength=n_rows), out=ptr[1:])
    return ptr, pairs[:, 1].astype(np.uint32)


def block_adjacency(
    edge_cell_l: NDArray[np.int32],
    edge_cell_r: NDArray[np.int32],
    n_cells: int,
    block_cells: int,
) -> tuple[NDArray[np.uint32], NDArray[np.uint32]]:
    """CSR of the other blocks each block shares an interior edge with (sorted, unique)."""
    n_blk = num_blocks(n_cells, block_cells)
    inner = np.asarray(edge_cell_r) >= 0
    bl = np.asarray(edge_cell_l[inner], dtype=np.int64) // block_cells
    br = np.asarray(edge_cell_r[inner], dtype=np.int64) // block_cells
    x = bl != br
    both = np.concatenate([np.column_stack([bl[x], br[x]]), np.column_stack([br[x], bl[x]])])
    return _csr_from_pairs(np.unique(both, axis=0), n_blk)


def block_two_ring(
    ptr: NDArray[np.uint32], idx: NDArray[np.uint32]
) -> tuple[NDArray[np.uint32], NDArray[np.uint32]]:
    """CSR of blocks within two adjacency steps (self excluded): one dilation over it is two."""
    n_blk = ptr.size - 1
    deg = np.diff(ptr).astype(np.int64)
    src1 = np.repeat(np.arange(n_blk), deg)
    dst1 = idx.astype(np.int64)
    deg2 = deg[dst1]
    src2 = np.repeat(src1, deg2)
    base = np.repeat(ptr[dst1].astype(np.int64), deg2)
    within = np.arange(int(deg2.sum())) - np.repeat(np.cumsum(deg2) - deg2, deg2)
    dst2 = idx[base + within].astype(np.int64)
    src = np.concatenate([src1, src2])
    dst = np.concatenate([dst1, dst2])
    keep = src != dst
    return _csr_from_pairs(np.unique(np.column_stack([src[keep], dst[keep]]), axis=0), n_blk)


def dilate_blocks(
    flags: NDArray[np.bool_], ptr: NDArray[np.uint32], idx: NDArray[np.uint32]
) -> NDArray[np.bool_]:
    """One ring of dilation: a block is set if it or any neighbour is set."""
    out = flags.copy()
    src = np.repeat(np.arange(flags.size), np.diff(ptr).astype(np.int64))
    np.logical_or.at(out, src, flags[idx])
    return out


__all__ = [
    "block_adjacency",
    "block_two_ring",
    "dilate_blocks",
    "edge_order_by_block",
    "edge_workgroup_block_ranges",
    "num_blocks",
    "remap_cell_edges",
]
