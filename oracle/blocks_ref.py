"""Block topology of the active set, for the FP32 oracle (TEST INFRASTRUCTURE).

The oracle restates the reference's resident kernels, which take three host
arrays besides the mesh (``inundation/mesh/blocks.py:47-118``): for each
256-edge workgroup of the block-sorted edges, the first and last block of
its edges' left cells; and the CSR of each block's two-ring (the blocks
within two adjacency steps), over which one dilation marks what two
dilations over the adjacency would. Here the block graph is a sparse
boolean matrix: A has a nonzero at (i, j) when an interior edge joins
blocks i != j, and the two-ring is the pattern of A + A A without its
diagonal. The CSR index arrays come out sorted by column, as the
reference's ``np.unique`` over (row, column) pairs gives them.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from numpy.typing import NDArray


def _pattern(m: sp.spmatrix) -> tuple[NDArray[np.uint32], NDArray[np.uint32]]:
    m = sp.csr_matrix(m, dtype=np.int64)
    m.setdiag(0)
    m.eliminate_zeros()
    m.sort_indices()
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32)


def block_adjacency(edge_cell_l, edge_cell_r, n_cells: int, block_cells: int):  # noqa: ANN001, ANN201
    """CSR (ptr, idx) of the blocks that share an interior edge with each block."""
    nb = -(-int(n_cells) // block_cells)
    cl = np.asarray(edge_cell_l, dtype=np.int64)
    cr = np.asarray(edge_cell_r, dtype=np.int64)
    inner = cr >= 0
    bi, bj = cl[inner] // block_cells, cr[inner] // block_cells
    ones = np.ones(bi.size, dtype=np.int64)
    a = sp.coo_matrix((ones, (bi, bj)), shape=(nb, nb))
    return _pattern((a + a.T) != 0)


def block_two_ring(ptr, idx):  # noqa: ANN001, ANN201
    """CSR of the blocks within two steps of each block, the block itself left out."""
    nb = np.asarray(ptr).size - 1
    a = sp.csr_matrix((np.ones(np.asarray(idx).size, dtype=np.int64), np.asarray(idx, dtype=np.int64),
                       np.asarray(ptr, dtype=np.int64)), shape=(nb, nb))
    return _pattern((a + a @ a) != 0)


def edge_workgroup_block_ranges(sorted_cell_l, workgroup: int, block_cells: int):  # noqa: ANN001, ANN201
    """(workgroups, 2): the block of the first and of the last edge of each run of ``workgroup`` edges."""
    blk = np.asarray(sorted_cell_l, dtype=np.int64) // block_cells
    if blk.size and (np.diff(blk) < 0).any():
        raise ValueError("edges must be sorted by the block of their left cell")
    starts = np.arange(0, blk.size, workgroup)
    ends = np.append(starts[1:], blk.size) - 1
    return np.stack([blk[starts], blk[ends]], axis=1).astype(np.uint32)


def dilate_blocks(flags, ptr, idx):  # noqa: ANN001, ANN201
    """One dilation: a block is set when it or a neighbour (over the CSR) is set."""
    f = np.asarray(flags, dtype=bool)
    nb = f.size
    a = sp.csr_matrix((np.ones(np.asarray(idx).size, dtype=np.int64), np.asarray(idx, dtype=np.int64),
                       np.asarray(ptr, dtype=np.int64)), shape=(nb, nb))
    return f | ((a @ f.astype(np.int64)) > 0)


__all__ = ["block_adjacency", "block_two_ring", "dilate_blocks", "edge_workgroup_block_ranges"]
