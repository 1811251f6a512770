"""One run script for the HIP backend and the FP32 oracle.

``drive`` feeds the same calls (parameters, state, sources, control block,
prologue, step blocks, retargets) to a backend context or to ``CpuRef`` and
returns everything they hold, so tests can compare the two bit for bit.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from inundation_amd.hip import lib as L
from inundation_amd.tuning import dt_to_ticks

FLAGS_DEFAULT = L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET


def setup(
    be: Any,
    n_cells: int,
    h,  # noqa: ANN001
    hu,  # noqa: ANN001
    hv,  # noqa: ANN001
    manning: float | np.ndarray,
    *,
    g: float = 9.81,
    dry_tol: float = 1e-3,
    cfl: float = 0.8,
    flags: int = FLAGS_DEFAULT,
    src_cells: np.ndarray | None = None,
    src_rates: np.ndarray | None = None,
) -> None:
    be.set_params(g, dry_tol, cfl, flags)
    n = np.broadcast_to(np.asarray(manning, dtype=np.float32), (n_cells,)).copy()
    be.set_manning(n)
    be.set_state(h, hu, hv, None)
    if src_cells is None:
        be.set_sources(np.empty(0, np.uint32), np.empty(0, np.float32))
    else:
        be.set_sources(src_cells, src_rates)


def start(be: Any, stop_s: float, dt_max: float, dt_init: float, cfl_safety: float = 1.0) -> Any:
    c = L.InuCtrl()
    c.now_ticks = 0
    c.stop_ticks = round(stop_s * 2**20)
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(dt_max)
    c.dry_dt_ticks = dt_to_ticks(min(dt_init, dt_max))
    c.cfl_safety = cfl_safety
    be.ctrl_write(c)
    be.prologue()
    return be.ctrl_read()


def snapshot(be: Any, n_cells: int, n_blocks: int) -> dict[str, np.ndarray]:
    out = {
        name: be.download(f, n_cells)
        for name, f in (("h", L.FIELD_H), ("hu", L.FIELD_HU), ("hv", L.FIELD_HV), ("h_lo", L.FIELD_H_LO))
    }
    out["wet"] = be.download(L.FIELD_BLOCK_WET, n_blocks, np.uint32)
    out["active"] = be.download(L.FIELD_BLOCK_ACTIVE, n_blocks, np.uint32)
    c = be.ctrl_read()
    out["ctrl"] = np.array([c.now_ticks, c.steps, c.dt_ticks], dtype=np.uint64)
    out["dt"] = be.dt_words()[:2].copy()
    return out


def assert_same(a: dict[str, np.ndarray], b: dict[str, np.ndarray], what: str = "") -> None:
    for k in a:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            same = np.array_equal(x.view(np.uint32), y.view(np.uint32))
        else:
            same = np.array_equal(x, y)
        if not same:
            if x.dtype == np.float32:
                d = np.flatnonzero(x.view(np.uint32) != y.view(np.uint32))
                raise AssertionError(
                    f"{what} {k}: {d.size} words differ; first {d[:5].tolist()} "
                    f"got {x[d[:5]].tolist()} want {y[d[:5]].tolist()}"
                )
            raise AssertionError(f"{what} {k}: got {x[:16]} want {y[:16]}")
