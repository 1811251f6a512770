"""The mesh slices the FP64 oracle runs on for large meshes (``tests/_slices.py``).

On a valley mesh, k FP64 Heun steps on a slice of core cells plus 2k rings
give the core exactly what the whole mesh gives; with fewer rings the cut
reaches the core.
"""

from __future__ import annotations

import numpy as np

from inundation_amd.hip.solver import device_mesh_arrays
from oracle import ref64
from tests._meshes import valley_grid, wet_lake
from tests._slices import ring_closure, submesh


def test_a_slice_with_2k_rings_steps_its_core_exactly() -> None:
    g = valley_grid(64, 32, 2.0)
    m = device_mesh_arrays(g)
    h, hu, hv = wet_lake(g)
    n = np.full(g.N, 0.03)
    k = 3
    full = ref64.run_dt_sequence(ref64.Mesh64.from_device_arrays(m), ref64.State64(h, hu, hv), n, [0.01] * k, 9.81,
                                 1e-3)
    core = np.arange(800, 900)
    for rings, exact in ((2 * k, True), (2 * k - 2, False)):
        sm, cells = submesh(m, ring_closure(m, core, rings))
        sub = ref64.run_dt_sequence(ref64.Mesh64.from_device_arrays(sm),
                                    ref64.State64(h[cells], hu[cells], hv[cells]), n[cells], [0.01] * k, 9.81, 1e-3)
        pos = np.searchsorted(cells, core)
        diff = max(np.abs(sub.h[pos] - full.h[core]).max(), np.abs(sub.hu[pos] - full.hu[core]).max())
        assert (diff == 0.0) == exact, (rings, diff)
