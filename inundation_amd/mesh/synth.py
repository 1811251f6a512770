"""Seeded synthetic meshes and terrains for the benchmark and test workloads.

- ``quad_grid``: axis-aligned quads (the lake-at-rest and valley cases);
- ``lattice_tri_quad``: a jittered lattice whose quads are split along their
  locally-Delaunay diagonal, a seeded fraction kept whole (the 16 M-cell
  rainfall mesh is ``lattice_tri_quad(3024, 8000.0, ...)``: 16 003 008 cells);
- ``voronoi_mesh``: Lloyd-relaxed Voronoi cells clipped to a box by mirroring;
- ``fbm_terrain``: spectral fractional Brownian motion sampled at centroids;
- ``gaussian_bumps``: a sum of Gaussian hills.

All generators take a ``numpy.random.Generator`` so every workload is fixed by
its seed.
"""

from __future__ import annotations

import numpy as np
from numpy.typing import NDArray


def quad_grid(nx: int, ny: int, dx: float) -> tuple[NDArray[np.float64], NDArray[np.int32]]:
    """``(verts, faces)`` of an ``nx`` by ``ny`` grid of ``dx`` quads, row-major, CCW."""
    xs, ys = np.meshgrid(np.arange(nx + 1) * dx, np.arange(ny + 1) * dx)
    verts = np.column_stack([xs.ravel(), ys.ravel()]).astype(np.float64)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v0 = (j * (nx + 1) + i).ravel()
    faces = np.column_stack([v0, v0 + 1, v0 + nx + 2, v0 + nx + 1]).astype(np.int32)
    return verts, faces


def _incircle(ax, ay, bx, by, cx, cy, dx, dy):  # noqa: ANN001, ANN202
    """Positive when d lies inside the circumcircle of the CCW triangle abc."""
    adx, ady = ax - dx, ay - dy
    bdx, bdy = bx - dx, by - dy
    cdx, cdy = cx - dx, cy - dy
    ad = adx * adx + ady * ady
    bd = bdx * bdx + bdy * bdy
    cd = cdx * cdx + cdy * cdy
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx)


def lattice_tri_quad(
    nq: int,
    length: float,
    rng: np.random.Generator,
    *,
    jitter: float = 0.2,
    quad_fraction: float = 0.25,
) -> tuple[NDArray[np.float64], NDArray[np.int32], NDArray[np.int32]]:
    """Mixed triangle/quad mesh of an ``nq`` x ``nq`` jittered lattice over a square.

    Interior vertices move by U[-jitter, jitter] lattice spacings on each axis.
    Exactly ``round(quad_fraction * nq^2)`` seeded quads stay whole; every
    other quad becomes two triangles along its locally-Delaunay diagonal.
    Returns ``(verts, faces_flat, face_offsets)``.
    """
    dx = length / nq
    n1 = nq + 1
    xs, ys = np.meshgrid(np.arange(n1) * dx, np.arange(n1) * dx)
    x = xs.ravel()
    y = ys.ravel()
    interior = np.zeros((n1, n1), dtype=bool)
    interior[1:-1, 1:-1] = True
    interior = interior.ravel()
    k = int(interior.sum())
    x[interior] += rng.uniform(-jitter, jitter, k) * dx
    y[interior] += rng.uniform(-jitter, jitter, k) * dx
    del xs, ys

    j, i = np.meshgrid(np.arange(nq), np.arange(nq), indexing="ij")
    a = (j * n1 + i).ravel().astype(np.int64)
    del j, i
    b = a + 1
    c = a + n1 + 1
    d = a + n1
    n_quads = nq * nq
    keep = np.zeros(n_quads, dtype=bool)
    keep[rng.permutation(n_quads)[: int(round(quad_fraction * n_quads))]] = True
    # Diagonal a-c unless d falls inside the circumcircle of abc (then b-d).
    flip = _incircle(x[a], y[a], x[b], y[b], x[c], y[c], x[d], y[d]) > 0.0

    n_tri_quads = n_quads - int(keep.sum())
    degree = np.where(keep, 4, 6)  # a split quad contributes 3 + 3 vertices
    start = np.zeros(n_quads + 1, dtype=np.int64)
    np.cumsum(degree, out=start[1:])
    faces = np.empty(int(start[-1]), dtype=np.int32)
    s = start[:-1]
    q = keep
    faces[s[q]] = a[q]
    faces[s[q] + 1] = b[q]
    faces[s[q] + 2] = c[q]
    faces[s[q] + 3] = d[q]
    t = ~keep & ~flip
    faces[s[t]] = a[t]
    faces[s[t] + 1] = b[t]
    faces[s[t] + 2] = c[t]
    faces[s[t] + 3] = a[t]
    faces[s[t] + 4] = c[t]
    faces[s[t] + 5] = d[t]
    t = ~keep & flip
    faces[s[t]] = a[t]
    faces[s[t] + 1] = b[t]
    faces[s[t] + 2] = d[t]
    faces[s[t] + 3] = b[t]
    faces[s[t] + 4] = c[t]
    faces[s[t] + 5] = d[t]
    # Offsets: one face per quad kept, two per split quad.
    n_cells = n_quads + n_tri_quads
    cell_deg = np.empty(n_cells, dtype=np.int64)
    first_cell = np.zeros(n_quads, dtype=np.int64)
    np.cumsum(np.where(keep, 1, 2)[:-1], out=first_cell[1:])
    cell_deg[first_cell[keep]] = 4
    cell_deg[first_cell[~keep]] = 3
    cell_deg[first_cell[~keep] + 1] = 3
    offsets = np.zeros(n_cells + 1, dtype=np.int64)
    np.cumsum(cell_deg, out=offsets[1:])
    if offsets[-1] >= 2**31:
        raise ValueError("mesh too large for int32 face offsets")
    verts = np.column_stack([x, y])
    return verts, faces, offsets.astype(np.int32)


def _mirrored(p: NDArray[np.float64], length: float, band: float) -> NDArray[np.float64]:
    """The points followed by their mirror images across each side of the box, for those within ``band`` of it."""
    parts = [p]
    for axis, side in ((0, 0.0), (0, length), (1, 0.0), (1, length)):
        near = p[np.abs(p[:, axis] - side) < band]
        m = near.copy()
        m[:, axis] = 2.0 * side - m[:, axis]
        parts.append(m)
    return np.concatenate(parts)


def _regions(pts: NDArray[np.float64], n_own: int) -> tuple[NDArray[np.float64], NDArray[np.int64], NDArray[np.int64]]:
    """Voronoi vertices of ``pts`` and the polygons (flat vertex ids, offsets) of the first ``n_own`` points."""
    from scipy.spatial import Voronoi

    vor = Voronoi(pts)
    regions = [vor.regions[r] for r in vor.point_region[:n_own].tolist()]
    deg = np.fromiter((len(r) for r in regions), dtype=np.int64, count=len(regions))
    offsets = np.zeros(deg.size + 1, dtype=np.int64)
    np.cumsum(deg, out=offsets[1:])
    flat = np.fromiter((v for r in regions for v in r), dtype=np.int64, count=int(offsets[-1]))
    if (flat < 0).any():
        raise RuntimeError("an unbounded Voronoi region inside the box")
    return vor.vertices, flat, offsets


def _centroids(pts: NDArray[np.float64], n_own: int) -> NDArray[np.float64]:
    """Area centroids of the Voronoi cells of the first ``n_own`` points of ``pts``."""
    verts, flat, offsets = _regions(pts, n_own)
    nxt = np.arange(1, flat.size + 1)
    nxt[offsets[1:] - 1] = offsets[:-1]  # each polygon's last vertex wraps to its first
    xs, ys = verts[flat, 0], verts[flat, 1]
    xn, yn = xs[nxt], ys[nxt]
    cr = xs * yn - xn * ys
    a6 = 3.0 * np.add.reduceat(cr, offsets[:-1])
    cx = np.add.reduceat((xs + xn) * cr, offsets[:-1]) / a6
    cy = np.add.reduceat((ys + yn) * cr, offsets[:-1]) / a6
    return np.column_stack([cx, cy])


def _tile_centroids(args: tuple) -> NDArray[np.float64]:  # noqa: ANN401
    own, halo = args
    return _centroids(np.concatenate([own, halo]), own.shape[0])


def _lloyd_step(pts: NDArray[np.float64], length: float, band: float, tiles: int, processes: int) -> NDArray[np.float64]:
    """One Lloyd step: every point to its cell's centroid.

    With ``tiles`` > 1 the box is cut into tiles x tiles squares; each tile's
    cells come from the Voronoi diagram of its own points plus every point
    (mirrors included) within a halo of 8 mean spacings around it, which
    holds all the neighbours a cell of uniform points has. The tiles run in
    ``processes`` worker processes (spawned, so a parent that already uses a
    GPU is safe).
    """
    everything = _mirrored(pts, length, band)
    if tiles <= 1:
        return _centroids(everything, pts.shape[0])
    halo_w = 2.0 * band
    edges = np.linspace(0.0, length, tiles + 1)
    ti = np.minimum((pts[:, 0] / length * tiles).astype(np.int64), tiles - 1)
    tj = np.minimum((pts[:, 1] / length * tiles).astype(np.int64), tiles - 1)
    jobs, owners = [], []
    for i in range(tiles):
        for j in range(tiles):
            mine = np.flatnonzero((ti == i) & (tj == j))
            x0, x1, y0, y1 = edges[i] - halo_w, edges[i + 1] + halo_w, edges[j] - halo_w, edges[j + 1] + halo_w
            inside = (everything[:, 0] >= x0) & (everything[:, 0] <= x1) & (everything[:, 1] >= y0) & (
                everything[:, 1] <= y1)
            inside[mine] = False
            jobs.append((pts[mine], everything[inside]))
            owners.append(mine)
    out = np.empty_like(pts)
    if processes > 1:
        import multiprocessing as mp

        with mp.get_context("spawn").Pool(min(processes, len(jobs))) as pool:
            results = pool.map(_tile_centroids, jobs)
    else:
        results = [_tile_centroids(job) for job in jobs]
    for mine, c in zip(owners, results):
        out[mine] = c
    return out


def voronoi_mesh(
    n_points: int, length: float, rng: np.random.Generator, *, lloyd: int = 2, processes: int = 1
) -> tuple[NDArray[np.float64], NDArray[np.int32], NDArray[np.int32]]:
    """Lloyd-relaxed Voronoi cells of seeded uniform points in ``[0, length]^2``.

    Points within a few mean spacings of a side are mirrored across it, so
    every cell is closed and the boundary cells end exactly on the square (the
    bisector of a point and its mirror is the side). The Lloyd steps of more
    than 400 000 points run on 4 x 4 tiles (the cells are those of the whole
    box; their centroids can differ from a whole-box step in the last bits),
    in ``processes`` worker processes; the result does not depend on
    ``processes``. Returns ``(verts, faces_flat, face_offsets)`` with the
    welded Voronoi vertices of the final diagram, which is always computed
    over the whole box.
    """
    pts = rng.uniform(0.0, length, size=(n_points, 2))
    band = 4.0 * length / np.sqrt(max(n_points, 1))
    tiles = 4 if n_points > 400_000 else 1
    for _ in range(lloyd):
        pts = np.clip(_lloyd_step(pts, length, band, tiles, processes), 1e-9 * length, length * (1.0 - 1e-9))
    verts_all, flat, offsets = _regions(_mirrored(pts, length, band), n_points)
    used = np.unique(flat)
    remap = np.full(verts_all.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    verts = np.clip(verts_all[used], 0.0, length)
    return verts, remap[flat].astype(np.int32), offsets.astype(np.int32)


def fbm_grid(m: int, hurst: float, rng: np.random.Generator) -> NDArray[np.float64]:
    """Periodic spectral fBm on an ``m`` x ``m`` grid, zero mean, unit rms.

    Amplitude ~ |k|^-(H + 1), i.e. power ~ |k|^-(2H + 2), random phases.
    """
    kx = np.fft.fftfreq(m)
    ky = np.fft.rfftfreq(m)
    kk = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    kk[0, 0] = 1.0
    amp = kk ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    spec = amp * np.exp(2j * np.pi * rng.random(amp.shape))
    field = np.fft.irfft2(spec, s=(m, m))
    field -= field.mean()
    return field / np.sqrt(np.mean(field * field))


def sample_bilinear(
    field: NDArray[np.float64], length: float, x: NDArray[np.float64], y: NDArray[np.float64]
) -> NDArray[np.float64]:
    """Periodic bilinear sample of a square grid covering ``[0, length]^2``."""
    m = field.shape[0]
    u = np.asarray(x, dtype=np.float64) / length * m
    v = np.asarray(y, dtype=np.float64) / length * m
    i0 = np.floor(u).astype(np.int64)
    j0 = np.floor(v).astype(np.int64)
    fu = u - i0
    fv = v - j0
    i0 %= m
    j0 %= m
    i1 = (i0 + 1) % m
    j1 = (j0 + 1) % m
    return (
        field[j0, i0] * (1 - fu) * (1 - fv)
        + field[j0, i1] * fu * (1 - fv)
        + field[j1, i0] * (1 - fu) * fv
        + field[j1, i1] * fu * fv
    )


def fbm_terrain(
    x: NDArray[np.float64],
    y: NDArray[np.float64],
    length: float,
    rng: np.random.Generator,
    *,
    hurst: float = 0.7,
    rms: float = 3.0,
    slope: float = 0.002,
    channel_width: float = 20.0,
    channel_depth: float = 1.5,
    grid: int = 1024,
) -> NDArray[np.float32]:
    """fBm bed (``rms`` metres) on a ``slope`` falling in +x, with a carved channel.

    The channel meanders along x around mid-height; cells within half its
    width of the centre line are lowered by ``channel_depth``.
    """
    z = rms * sample_bilinear(fbm_grid(grid, hurst, rng), length, x, y) - slope * x
    centre = 0.5 * length + 0.1 * length * np.sin(2.0 * np.pi * x / length)
    z = np.where(np.abs(y - centre) <= 0.5 * channel_width, z - channel_depth, z)
    return z.astype(np.float32)


def gaussian_bumps(
    x: NDArray[np.float64],
    y: NDArray[np.float64],
    length: float,
    rng: np.random.Generator,
    *,
    count: int = 12,
    sigma: tuple[float, float] = (30.0, 80.0),
    amplitude: tuple[float, float] = (0.2, 1.2),
) -> NDArray[np.float32]:
    """Sum of ``count`` Gaussian bumps with seeded centres, widths and heights."""
    cx = rng.uniform(0.0, length, count)
    cy = rng.uniform(0.0, length, count)
    sg = rng.uniform(*sigma, count)
    am = rng.uniform(*amplitude, count)
    z = np.zeros_like(np.asarray(x, dtype=np.float64))
    for k in range(count):
        z += am[k] * np.exp(-((x - cx[k]) ** 2 + (y - cy[k]) ** 2) / (2.0 * sg[k] ** 2))
    return z.astype(np.float32)


__all__ = [
    "fbm_grid",
    "fbm_terrain",
    "gaussian_bumps",
    "lattice_tri_quad",
    "quad_grid",
    "sample_bilinear",
    "voronoi_mesh",
]
