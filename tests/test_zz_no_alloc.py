"""inu_set_source_rates_noalloc against inu_set_source_rates (needs an MI355X; collected last).

Two contexts on the same mesh, sources and state; one takes each phase's
rates through the allocating entry point, the other through the one that
stages them in the library's fixed device array. After each phase's steps
the two states are the same bits. The second call is smaller than the
first, so stale entries left in the shared array must not be read. One
scenario and a two-scenario batch (whose cell ids the call converts).
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays
from tests import _drive as D
from tests._meshes import MANNING, valley_grid
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu


def _contexts(n_scen: int):  # noqa: ANN202
    grid = valley_grid(128, 64, 2.0)
    m = device_mesh_arrays(grid)
    ctxs = [gpu_context(), gpu_context()]
    for c in ctxs:
        if n_scen == 1:
            c.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
        else:
            c.set_batch([m] * n_scen, [0] * n_scen)
    return grid, ctxs


def _state(grid, n_scen: int):  # noqa: ANN001, ANN202
    x = np.tile(grid.centroid[:, 0], n_scen)
    h = np.where(x < 60.0, 0.5, 0.0).astype(np.float32)
    z = np.zeros_like(h)
    return h, z, z.copy()


@pytest.mark.parametrize("n_scen", [1, 2])
def test_noalloc_rates_match_the_allocating_call(n_scen: int) -> None:
    grid, (ref, noal) = _contexts(n_scen)
    n = grid.N * n_scen
    try:
        cells = np.arange(n, dtype=np.uint32)
        rain = np.full(n, 2e-5, np.float32)
        h, hu, hv = _state(grid, n_scen)
        for c in (ref, noal):
            D.setup(c, n, h, hu, hv, MANNING, dry_tol=1e-3, src_cells=cells, src_rates=rain)
            c.ctrl_write([_ctrl() for _ in range(n_scen)] if n_scen > 1 else _ctrl())
            c.prologue()
        rng = np.random.default_rng(11)
        for phase, count in enumerate((n, 300)):  # a smaller second call
            sel = np.sort(rng.choice(n, count, replace=False)).astype(np.uint32)
            rates = rng.uniform(0.0, 5e-4, count).astype(np.float32)
            ref.set_source_rates(sel, rates)
            noal.set_source_rates_noalloc(sel, rates)
            for c in (ref, noal):
                c.steps(12, use_graph=True)
            for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO):
                a, b = ref.download(f, n), noal.download(f, n)
                d = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
                assert d.size == 0, f"phase {phase} field {f}: {d.size} words differ"
            ca = ref.ctrls_read()
            cb = noal.ctrls_read()
            assert [bytes(x) for x in ca] == [bytes(x) for x in cb]
            assert all(int(x.steps) == 12 * (phase + 1) for x in ca)
        with pytest.raises(ValueError, match="at most 65536"):
            big = np.arange(65537, dtype=np.uint32) % n
            noal.set_source_rates_noalloc(big, np.zeros(big.size, np.float32))
    finally:
        ref.close()
        noal.close()


def _ctrl() -> L.InuCtrl:
    from inundation_amd.tuning import dt_to_ticks

    c = L.InuCtrl()
    c.now_ticks = 0
    c.stop_ticks = 2**40
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(0.25)
    c.dry_dt_ticks = dt_to_ticks(0.05)
    c.cfl_safety = 1.0
    return c
