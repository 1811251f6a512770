"""Write the golden fixtures of ``tests/golden/`` from the reference's own code.

Run in a checkout that has the reference at ``/root/reference`` (read only):

    python tests/golden/make_golden.py

It imports the reference's pure-NumPy geometry build, Hilbert ordering, block
helpers and its FP64 test oracle (``tests/unit/_cpu_step.py``) and stores
inputs and outputs as ``.npz`` data:

- ``geometry_*.npz``: every ``MeshGeometry`` array after ``build_geometry`` and
  after ``hilbert_reorder`` (cell, edge and CSR numbering, FP32 rounding) for
  a few small meshes, including clockwise and mixed-degree cells;
- ``blocks_*.npz``: the active-set block helpers on those meshes;
- ``cpu_step_valley.npz``: 8-step FP64 trajectories of the four step-parity
  cases (reference ``test_gpu_step_parity.py:81-115``) on a 32 x 16 valley,
  and the FP64 CFL bound;
- ``hllc_vectors.npz``: HLLC fluxes and speeds of seeded random face states;
- ``artifact_v6_patch.npz``: a geometry-cache artifact written by the
  reference's ``build_geometry_cache`` from a small OBJ mesh, and
  ``mesh_obj_cache.npz``: that OBJ's text, the reference's loader output
  and its cache keys (``python tests/golden/make_golden.py cache`` writes
  only these two).

Only data is written; nothing of the reference's source is copied.
"""

from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def _ref_modules():  # noqa: ANN202
    sys.path.insert(0, str(REF))
    import inundation.mesh.blocks as blocks
    import inundation.mesh.geometry as geometry

    spec = importlib.util.spec_from_file_location("_ref_cpu_step", REF / "tests" / "unit" / "_cpu_step.py")
    cpu_step = importlib.util.module_from_spec(spec)
    sys.modules["_ref_cpu_step"] = cpu_step
    spec.loader.exec_module(cpu_step)
    return geometry, blocks, cpu_step


GEOM_FIELDS = (
    "origin", "centroid", "area", "zb", "cell_vertices", "cell_vertex_ptr", "cell_bbox", "edge_len",
    "edge_nx", "edge_ny", "edge_dzL", "edge_dzR", "edge_cellL", "edge_cellR", "edge_sideL",
    "edge_sideR", "cell_edge_ptr", "cell_edge_idx", "cell_nbr_idx", "cell_edge_side", "cell_edge_idx_signed",
)


def meshes() -> dict[str, tuple]:
    """(verts, faces, offsets or None, zb, manning or None) of each fixture mesh."""
    sys.path.insert(0, str(OUT.parent.parent))
    from inundation_amd.mesh.synth import lattice_tri_quad, quad_grid

    out = {}
    # 2 x 2 quads, as the reference cache self-check builds it (cache.py:305-321).
    v, f = quad_grid(2, 2, 1.0)
    out["quads2x2"] = (v, f, None, np.array([0.0, 0.1, 0.2, 0.3], np.float32), None)
    # 4 x 4 vertices, two triangles per square (test_inundation_signed_csr.py:19-35).
    xs, ys = np.meshgrid(np.arange(4, dtype=np.float32), np.arange(4, dtype=np.float32))
    v = np.stack([xs.ravel(), ys.ravel()], axis=1)
    tri = []
    for j in range(3):
        for i in range(3):
            a, b, c, d = j * 4 + i, j * 4 + i + 1, (j + 1) * 4 + i + 1, (j + 1) * 4 + i
            tri += [[a, b, c], [a, c, d]]
    out["tri4x4"] = (v, np.array(tri, np.int32), None, np.linspace(0, 1, 18).astype(np.float32), None)
    # A 32 x 16 valley of 2 m quads at a 300 m datum.
    v, f = quad_grid(32, 16, 2.0)
    cx = (np.arange(32 * 16) % 32 + 0.5) * 2.0
    cy = (np.arange(32 * 16) // 32 + 0.5) * 2.0
    zb = (300.0 - 0.002 * cx - 0.3 * np.exp(-(((cy - 16.0) / 6.4) ** 2))).astype(np.float32)
    out["valley32x16"] = (v.astype(np.float32), f, None, zb, None)
    # Mixed triangles and quads on a jittered lattice, some cells clockwise,
    # plus pentagons made by merging triangle pairs is not needed: a hexagon
    # fan supplies the higher degrees below.
    rng = np.random.default_rng(7)
    v, f, o = lattice_tri_quad(9, 9.0, rng)
    deg = np.diff(o)
    f = f.copy()
    for c in range(0, deg.size, 3):  # reverse every third cell
        f[o[c] : o[c + 1]] = f[o[c] : o[c + 1]][::-1]
    zb = rng.uniform(0, 1, deg.size).astype(np.float32)
    n = rng.uniform(0.02, 0.05, deg.size).astype(np.float32)
    out["mixed9"] = (v, f, o, zb, n)
    # A hexagon ring: degrees 3 to 6 around a centre.
    ang = np.linspace(0, 2 * np.pi, 7)[:-1]
    ring = np.column_stack([np.cos(ang), np.sin(ang)]) * 2.0
    outer = np.column_stack([np.cos(ang + np.pi / 6), np.sin(ang + np.pi / 6)]) * 4.0
    v = np.vstack([ring, outer, [[0.0, 0.0]]])
    faces = [[0, 1, 2, 3, 4, 5]]
    for k in range(6):
        faces.append([k, 6 + k, (k + 1) % 6])
        faces.append([(k + 1) % 6, 6 + k, 6 + (k + 1) % 6])
    flat = np.concatenate([np.asarray(x, np.int32) for x in faces])
    offs = np.r_[0, np.cumsum([len(x) for x in faces])].astype(np.int32)
    out["hexring"] = (v, flat, offs, np.arange(len(faces), dtype=np.float32) * 0.1, None)
    return out


def geometry_fixtures(geometry, blocks) -> None:  # noqa: ANN001
    for name, (v, f, o, zb, n) in meshes().items():
        g = geometry.build_geometry(v, f, o, zb_from_file=zb, manning_n_from_file=n)
        r, perm = geometry.hilbert_reorder(g, verbose=False)
        data = {"in_verts": np.asarray(v), "in_faces": np.asarray(f), "in_zb": zb}
        if o is not None:
            data["in_offsets"] = np.asarray(o)
        if n is not None:
            data["in_manning"] = n
        for tag, geom in (("built", g), ("reordered", r)):
            for k in GEOM_FIELDS:
                data[f"{tag}_{k}"] = np.asarray(getattr(geom, k))
            data[f"{tag}_counts"] = np.array([geom.N, geom.E, geom.V, geom.max_degree], np.int64)
        data["perm"] = perm
        # Block helpers with small blocks so these meshes have several.
        for bc in (2, 4):
            order = blocks.edge_order_by_block(r.edge_cellL, bc)
            data[f"b{bc}_order"] = order
            data[f"b{bc}_remap"] = blocks.remap_cell_edges(r.cell_edge_idx_signed, order)
            data[f"b{bc}_ranges"] = blocks.edge_workgroup_block_ranges(r.edge_cellL[order], 4, bc)
            ptr, idx = blocks.block_adjacency(r.edge_cellL, r.edge_cellR, r.N, bc)
            p2, i2 = blocks.block_two_ring(ptr, idx)
            data[f"b{bc}_adj_ptr"], data[f"b{bc}_adj_idx"] = ptr, idx
            data[f"b{bc}_ring_ptr"], data[f"b{bc}_ring_idx"] = p2, i2
        np.savez_compressed(OUT / f"geometry_{name}.npz", **data)
        print(name, g.N, g.E)


def cpu_step_fixtures(geometry, cpu_step) -> None:  # noqa: ANN001
    v, f, _, zb, _ = meshes()["valley32x16"]
    g, _ = geometry.hilbert_reorder(geometry.build_geometry(v, f, zb_from_file=zb), verbose=False)
    x = g.centroid[:, 0].astype(np.float64)
    y = g.centroid[:, 1].astype(np.float64)
    lx, ly = 64.0, 32.0
    dam = np.where((x < lx * 0.25) & (np.abs(y - ly / 2) < ly * 0.15), 1.0, 0.0)
    sheet = np.where(x < lx * 0.3, 0.02, 0.0)
    lake = float(g.zb.max()) + 0.5 - g.zb.astype(np.float64)
    cases = {
        "wet_lake": (lake, 0.1 * np.sin(2 * np.pi * x / lx) * lake, np.zeros_like(x)),
        "dam_break": (dam, np.zeros_like(x), np.zeros_like(x)),
        "draining": (sheet, np.where(sheet > 0, 0.05, 0.0), np.zeros_like(x)),
        "dry_momentum": (dam, np.where(x > lx * 0.8, 0.1, 0.0), -np.where(x > lx * 0.8, 0.1, 0.0)),
    }
    data = {}
    n = np.full(g.N, 0.03, np.float32)
    for k, (h, hu, hv) in cases.items():
        h32, hu32, hv32 = (np.asarray(a, np.float32) for a in (h, hu, hv))
        data[f"{k}_h0"], data[f"{k}_hu0"], data[f"{k}_hv0"] = h32, hu32, hv32
        s = cpu_step.run_steps(g, h32, hu32, hv32, n, 2.0**-10, 8, g=9.81, dry_tol=1e-3)
        data[f"{k}_h8"], data[f"{k}_hu8"], data[f"{k}_hv8"] = s.h, s.hu, s.hv
        st = cpu_step.State(h32.astype(np.float64), hu32.astype(np.float64), hv32.astype(np.float64))
        data[f"{k}_cfl"] = np.array(cpu_step.cfl_dt(g, st, g=9.81, dry_tol=1e-3, cfl=0.8))
    np.savez_compressed(OUT / "cpu_step_valley.npz", **data)
    print("cpu_step", g.N)


def hllc_fixtures(cpu_step) -> None:  # noqa: ANN001
    rng = np.random.default_rng(11)
    m = 1000
    dry = 1e-3
    depth = np.concatenate([rng.uniform(0, 3 * dry, m // 2), 10.0 ** rng.uniform(-3, 1, m // 2)])
    hL = rng.permutation(depth)
    hR = rng.permutation(depth)
    hL[:200] = 0.0
    hR[200:400] = 0.0
    qnL, qtL, qnR, qtR = (rng.normal(0, 0.5, m) * np.maximum(h, 0) for h in (hL, hL, hR, hR))
    Fh, Fhn, Fht, smax = cpu_step.hllc(hL, qnL, qtL, hR, qnR, qtR, g=9.81, dry_tol=dry)
    unL, unR = rng.normal(0, 1, m), rng.normal(0, 1, m)
    SL, SR = cpu_step.hllc_speeds(hL, unL, hR, unR, g=9.81, dry_tol=dry)
    np.savez_compressed(
        OUT / "hllc_vectors.npz",
        hL=hL, qnL=qnL, qtL=qtL, hR=hR, qnR=qnR, qtR=qtR, Fh=Fh, Fhn=Fhn, Fht=Fht, smax=smax,
        unL=unL, unR=unR, SL=SL, SR=SR, dry=np.array(dry), g=np.array(9.81),
    )
    print("hllc", m)


# A 3 x 2 patch of quads and triangles with vertex elevations, one face
# clockwise, one vertex listed but unused, "v/vt" face tokens.
OBJ_TEXT = """\
# mixed patch
v 0 0 1.0
v 1 0 1.5
v 2 0 2.0
v 3 0 2.5
v 0 1 1.2
v 1 1 1.7
v 2 1 2.2
v 3 1 2.7
v 0 2 1.4
v 1 2 1.9
v 2 2 2.4
v 3 2 2.9
v 9 9 9.0
f 1 2 6 5
f 2/1 3/2 7/3
f 2 7 6
f 3 4 8 7
f 5 6 10 9
f 6 10 11
f 6 7 11
f 7 8 12 11
"""


def cache_fixtures() -> None:
    import tempfile

    sys.path.insert(0, str(REF))
    import inundation.mesh.cache as cache
    import inundation.mesh.loader as loader

    with tempfile.TemporaryDirectory() as tmp:
        obj = Path(tmp) / "patch.obj"
        obj.write_text(OBJ_TEXT, encoding="ascii")
        art = cache.build_geometry_cache(obj, Path(tmp) / "patch.geometry.npz", use_hilbert_reorder=True)
        (OUT / "artifact_v6_patch.npz").write_bytes(Path(art).read_bytes())
        v, f, o, zb, n = loader.load_mesh_file(str(obj))
        assert n is None
        np.savez_compressed(
            OUT / "mesh_obj_cache.npz",
            obj_text=np.frombuffer(OBJ_TEXT.encode("ascii"), dtype=np.uint8),
            verts=v, faces=f, offsets=o, zb=zb,
            key_hilbert=np.array(cache.geometry_cache_key(obj, use_hilbert_reorder=True)),
            key_plain=np.array(cache.geometry_cache_key(obj, use_hilbert_reorder=False)),
        )
    print("cache", cache.GEOMETRY_CACHE_VERSION)


def main() -> None:
    if sys.argv[1:] == ["cache"]:
        cache_fixtures()
        return
    geometry, blocks, cpu_step = _ref_modules()
    geometry_fixtures(geometry, blocks)
    cpu_step_fixtures(geometry, cpu_step)
    hllc_fixtures(cpu_step)
    cache_fixtures()


if __name__ == "__main__":
    main()
