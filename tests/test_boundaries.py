"""Stretch picking and the conveyance split (``inundation_amd/mesh/boundaries.py``) on random sections.

The split is checked against what fixes it: every wetted edge of a section
stands at one water level, the edges left dry lie at or above that level,
and the section carries the discharge; the polyline distance against a
point-by-point projection onto each segment.
"""

from __future__ import annotations

import numpy as np

from inundation_amd.mesh.boundaries import conveyance_split, distance_to_line


def test_random_sections_share_one_level_and_carry_the_discharge() -> None:
    rng = np.random.default_rng(0)
    for _ in range(300):
        n = int(rng.integers(1, 12))
        length = rng.uniform(1, 20, n).astype(np.float32)
        bed = rng.uniform(0, 3, n).astype(np.float32)
        manning = rng.uniform(0.02, 0.08, n).astype(np.float32)
        q_total, slope = float(rng.uniform(0.01, 500)), float(rng.uniform(1e-4, 1e-2))
        q = conveyance_split(q_total, length, bed, manning, slope)
        assert abs(float((q.astype(np.float64) * length).sum()) - q_total) <= 1e-6 * q_total
        k = np.sqrt(slope) / manning.astype(np.float64)
        level = bed + (np.maximum(q, 0.0) / k) ** 0.6
        wet = q > 0
        assert wet.any() and np.ptp(level[wet]) < 1e-4 * max(1.0, float(np.abs(level[wet]).max()))
        if (~wet).any():
            assert level[wet].max() <= bed[~wet].min() + 1e-5


def test_polyline_distance_is_the_nearest_segment_projection() -> None:
    rng = np.random.default_rng(1)
    line = rng.uniform(0, 100, (5, 2))
    pts = rng.uniform(-20, 120, (400, 2))
    got = distance_to_line(pts, line)
    want = np.full(pts.shape[0], np.inf)
    for a, b in zip(line[:-1], line[1:]):
        for i, p in enumerate(pts):
            t = min(max(float(np.dot(p - a, b - a) / np.dot(b - a, b - a)), 0.0), 1.0)
            want[i] = min(want[i], float(np.linalg.norm(p - (a + t * (b - a)))))
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
