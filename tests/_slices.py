"""A slice of a large mesh for the FP64 oracle: core cells plus the rings their result depends on.

One Heun step reads each cell's neighbours twice (once per stage), so after k
steps a cell depends on the cells within 2k edges of it. ``submesh`` keeps
the core cells and ``rings`` rings around them; edges leading out of the kept
cells become walls. The FP64 restatement run on the slice gives, for the core
cells, the same steps as on the whole mesh while ``rings >= 2k``; only the
outer rings feel the cut.
"""

from __future__ import annotations

from typing import Any

import numpy as np
from numpy.typing import NDArray


def ring_closure(m: dict[str, Any], core: NDArray[np.int64], rings: int) -> NDArray[np.bool_]:
    """Mask of the cells within ``rings`` edges of ``core``."""
    n = int(m["area"].size)
    cl = np.asarray(m["cell_l"], dtype=np.int64)
    cr = np.asarray(m["cell_r"], dtype=np.int64)
    inner = cr >= 0
    a, b = cl[inner], cr[inner]
    keep = np.zeros(n, dtype=bool)
    keep[core] = True
    for _ in range(rings):
        grow = keep.copy()
        grow[b[keep[a]]] = True
        grow[a[keep[b]]] = True
        keep = grow
    return keep


def submesh(m: dict[str, Any], keep: NDArray[np.bool_]) -> tuple[dict[str, Any], NDArray[np.int64]]:
    """Device-order mesh arrays of the kept cells (cells renumbered in order) and the kept cell ids.

    An edge with one side outside the kept cells becomes a wall of the inside
    cell: its left side is that cell (normal and bed steps turned round when
    it was the right side), and its CSR slot is a left slot.
    """
    cells = np.flatnonzero(keep).astype(np.int64)
    new_id = np.full(keep.size, -1, dtype=np.int64)
    new_id[cells] = np.arange(cells.size)
    cl = np.asarray(m["cell_l"], dtype=np.int64)
    cr = np.asarray(m["cell_r"], dtype=np.int64)
    in_l = keep[cl]
    in_r = (cr >= 0) & keep[np.maximum(cr, 0)]
    edges = np.flatnonzero(in_l | in_r)
    flip = ~in_l[edges]  # only the right cell is kept
    dz = np.asarray(m["dz2"], dtype=np.float32).reshape(-1, 2)[edges]
    nx = np.asarray(m["nx"], dtype=np.float32)[edges]
    ny = np.asarray(m["ny"], dtype=np.float32)[edges]
    left = np.where(flip, cr[edges], cl[edges])
    right = np.where(flip | ~in_r[edges], -1, cr[edges])
    sub_e = np.full(cl.size, -1, dtype=np.int64)
    sub_e[edges] = np.arange(edges.size)
    out = {
        "cell_l": new_id[left].astype(np.int32),
        "cell_r": np.where(right >= 0, new_id[np.maximum(right, 0)], -1).astype(np.int32),
        "nx": np.where(flip, -nx, nx).astype(np.float32),
        "ny": np.where(flip, -ny, ny).astype(np.float32),
        "len": np.asarray(m["len"], dtype=np.float32)[edges],
        "dz2": np.column_stack([np.where(flip, dz[:, 1], dz[:, 0]), np.where(flip, dz[:, 0], dz[:, 1])])
        .astype(np.float32)
        .ravel(),
        "area": np.asarray(m["area"], dtype=np.float32)[cells],
    }
    cptr = np.asarray(m["cptr"], dtype=np.int64)
    cidx = np.asarray(m["cidx"], dtype=np.int64)
    deg = cptr[cells + 1] - cptr[cells]
    slots = np.concatenate([np.arange(cptr[c], cptr[c + 1]) for c in cells.tolist()]) if cells.size else np.zeros(0, np.int64)
    e_old = cidx[slots] >> 1
    right_slot = (cidx[slots] & 1).astype(bool)
    e_new = sub_e[e_old]
    assert (e_new >= 0).all()
    right_slot &= ~flip[e_new]
    out["cptr"] = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    out["cidx"] = ((e_new << 1) | right_slot.astype(np.int64)).astype(np.uint32)
    return out, cells
