"""Every tile level of the stage kernels, and the autotuner's choice, bit for bit against the FP32 oracle.

A tile level is how many edge records and CSR slots each thread loads
before the stage kernels' first barrier ((2, 4), (3, 6), (4, 8)); the rest
are loaded after it. The mesh is a Voronoi mesh with every side split in
two, so every block holds more than 1024 local edges and 2048 CSR slots and
the loops after the barrier run at every level. Both stages are compared
(the predictor state too), with and without sources, compensated and plain.
"""

from __future__ import annotations

import numpy as np
import pytest

from inundation_amd.hip import lib as L
from inundation_amd.hip.solver import device_mesh_arrays
from oracle import cpu_ref
from tests import _drive as D
from tests._meshes import dense_voronoi
from tests.conftest import gpu_context

pytestmark = pytest.mark.gpu

PRED = (("h1", L.FIELD_H1), ("hu1", L.FIELD_HU1), ("hv1", L.FIELD_HV1), ("h1_lo", L.FIELD_H1_LO))


@pytest.fixture(scope="module")
def ctx():  # noqa: ANN201
    c = gpu_context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def geom():  # noqa: ANN201
    return dense_voronoi()


@pytest.fixture(scope="module")
def mesh(geom):  # noqa: ANN001, ANN201
    m = device_mesh_arrays(geom)
    cl = m["cell_l"]
    cr = m["cell_r"]
    nb = -(-geom.N // 256)
    own = np.bincount(cl // 256, minlength=nb)
    halo = np.bincount(cr[(cr >= 0) & (cr // 256 != cl // 256)] // 256, minlength=nb)
    slots = np.add.reduceat(np.diff(m["cptr"]).astype(np.int64), np.arange(0, geom.N, 256))
    # Full blocks must exceed the largest level's preload.
    assert (own + halo)[:-1].min() > 4 * 256 and slots[:-1].min() > 8 * 256
    return m


def _state(geom):  # noqa: ANN001, ANN202
    x, y = geom.centroid[:, 0].astype(np.float64), geom.centroid[:, 1].astype(np.float64)
    lx, ly = float(x.max()), float(y.max())
    pool = (x - 0.3 * lx) ** 2 + (y - 0.5 * ly) ** 2 < (0.2 * lx) ** 2
    eta = float(geom.zb[pool].max()) + 0.5
    h = np.where(pool, np.maximum(eta - geom.zb, 0.0), 0.0).astype(np.float32)
    z = np.zeros(geom.N, np.float32)
    return h, z, z.copy()


def _sources(geom):  # noqa: ANN001, ANN202
    """Rain of 60 mm/h on the right third, plus a point inflow there."""
    x = geom.centroid[:, 0]
    cells = np.flatnonzero(x > 0.66 * float(x.max())).astype(np.uint32)
    rates = np.full(cells.size, 60.0 / 1000.0 / 3600.0, np.float32)
    rates[cells.size // 2] += np.float32(2.0 / geom.area[cells[cells.size // 2]])
    return cells, rates


def _snap(be, geom, listed=None) -> dict[str, np.ndarray]:  # noqa: ANN001
    """The step state; with ``listed`` (the block flags the last step ran with), also the predictor
    state of those blocks (skipped blocks keep stale predictor words that nothing reads)."""
    s = D.snapshot(be, geom.N, -(-geom.N // 256))
    if listed is not None:
        cells = np.repeat(listed != 0, 256)[: geom.N]
        for name, f in PRED:
            s[name] = be.download(f, geom.N)[cells]
    return s


FLAGS = {"comp+active": D.FLAGS_DEFAULT, "plain+all": L.FLAG_DETERMINISTIC}


@pytest.mark.parametrize("sources", [False, True], ids=["no-src", "src"])
@pytest.mark.parametrize("flags", list(FLAGS.values()), ids=list(FLAGS))
@pytest.mark.parametrize("tile", [0, 1, 2, "auto"])
def test_tile_level_is_bitwise_the_fp32_oracle(ctx, geom, mesh, tile, flags, sources) -> None:  # noqa: ANN001
    print(f"tile {tile} flags {flags} sources {sources}", flush=True)
    ctx.set_mesh(mesh["cell_l"], mesh["cell_r"], mesh["nx"], mesh["ny"], mesh["len"], mesh["dz2"], mesh["area"],
                 mesh["cptr"], mesh["cidx"])
    ref = cpu_ref.CpuRef(mesh)
    h, hu, hv = _state(geom)
    cells, rates = _sources(geom) if sources else (None, None)
    for be in (ctx, ref):
        D.setup(be, geom.N, h, hu, hv, 0.03, dry_tol=1e-4, flags=flags, src_cells=cells, src_rates=rates)
        D.start(be, 60.0, 1.0, 0.05)
    if tile == "auto":
        choice = ctx.autotune(3)
        assert choice["tile"] in (0, 1, 2) and min(choice["us_per_launch"]) > 0
        # the autotuner leaves the state as it was: still the oracle's
        D.assert_same(_snap(ctx, geom), _snap(ref, geom), "after autotune")
    else:
        ctx.set_tile(tile)
    for k, n in enumerate((1, 4, 14)):
        ctx.steps(n, use_graph=k > 0)
        ref.steps(n)
        D.assert_same(_snap(ctx, geom), _snap(ref, geom), f"tile {tile}: after block {k}")
    listed = ref.download(L.FIELD_BLOCK_ACTIVE, -(-geom.N // 256), np.uint32)
    assert listed.any()
    ctx.steps(1, use_graph=False)
    ref.steps(1)
    D.assert_same(_snap(ctx, geom, listed), _snap(ref, geom, listed), f"tile {tile}: predictor and corrector")
    assert int(ctx.ctrl_read().steps) == 20
