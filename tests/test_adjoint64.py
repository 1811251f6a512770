"""The FP64 reference gradients (``oracle/adjoint64.py``) pinned.

- its forward is the FP64 oracle ``oracle/ref64.py`` (itself pinned to the
  reference's FP64 test oracle) given the same bed steps, to 1e-12;
- its gradients with respect to z and n agree with FP64 central differences
  on an all-wet case with no bed ties (no kink), to 1e-4 relative (the
  differences' own round-off on the small Manning gradients);
- a planted mistake (Manning gradient of the wrong power of h) fails the
  finite-difference check.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from inundation_amd.hip.solver import device_mesh_arrays
from oracle import adjoint64 as A
from oracle import ref64
from tests._meshes import valley_grid


@pytest.fixture(scope="module")
def case():  # noqa: ANN201
    g = valley_grid(16, 8, 2.0)
    m = device_mesh_arrays(g)
    x = g.centroid[:, 0].astype(np.float64)
    y = g.centroid[:, 1].astype(np.float64)
    # The valley is symmetric about its centre line; a small perturbation
    # removes the bed ties there (a tie is a kink of the bed steps).
    z = g.zb.astype(np.float64) + 1e-3 * np.sin(3.7 * x + 1.3 * y)
    h = np.maximum(float(z.max()) + 0.4 - z, 0.0)
    hu = 0.15 * np.sin(2 * np.pi * x / 32.0) * h
    hv = 0.05 * np.cos(2 * np.pi * x / 32.0) * h
    n = np.linspace(0.02, 0.06, g.N)
    dts = np.full(12, 0.05)
    gauges = np.array([3, 40, 77, 120])
    return g, m, z, n, (h, hu, hv), dts, gauges


def test_forward_is_the_fp64_oracle(case) -> None:  # noqa: ANN001
    g, m, z, n, s0, dts, _ = case
    mesh = A.MeshT.from_device_arrays(m)
    dzl, dzr = A.bed_steps(mesh, torch.as_tensor(z))
    m64 = ref64.Mesh64.from_device_arrays(m)
    m64.dzl, m64.dzr = dzl.numpy(), dzr.numpy()
    want = ref64.run_dt_sequence(m64, ref64.State64(*[a.copy() for a in s0]), n, dts, 9.81, 1e-4)
    h, hu, hv = (torch.as_tensor(a) for a in s0)
    for dt in dts.tolist():
        h, hu, hv = A.heun_step(mesh, h, hu, hv, torch.as_tensor(n), dzl, dzr, dt, 9.81, 1e-4)
    np.testing.assert_allclose(h.numpy(), want.h, rtol=0, atol=1e-12)
    np.testing.assert_allclose(hu.numpy(), want.hu, rtol=0, atol=1e-12)
    np.testing.assert_allclose(hv.numpy(), want.hv, rtol=0, atol=1e-12)


def _loss(m, z, n, s0, dts, gauges, obs):  # noqa: ANN001, ANN202
    series = A.gauge_series(m, z, n, s0, dts, None, gauges, 3)
    return float(((series - obs) ** 2).sum())


def _fd_check(case, grads_fn) -> float:  # noqa: ANN001
    g, m, z, n, s0, dts, gauges = case
    obs = A.gauge_series(m, z, n * 1.5, s0, dts, None, gauges, 3) - 0.01
    _, gz, gn = grads_fn(m, z, n, s0, dts, None, gauges, obs, 3)
    rng = np.random.default_rng(0)
    worst = 0.0
    for i in rng.choice(g.N, 6, replace=False).tolist():
        for vec, grad, eps in ((z, gz, 1e-6), (n, gn, 1e-5)):
            up, dn = vec.copy(), vec.copy()
            up[i] += eps
            dn[i] -= eps
            args = (up, n) if vec is z else (z, up)
            argd = (dn, n) if vec is z else (z, dn)
            fd = (_loss(m, *args, s0, dts, gauges, obs) - _loss(m, *argd, s0, dts, gauges, obs)) / (2 * eps)
            scale = max(abs(fd), 1e-3 * np.abs(grad).max())
            worst = max(worst, abs(grad[i] - fd) / scale)
    return worst


def test_gradients_match_central_differences(case) -> None:  # noqa: ANN001
    # Central differences in FP64 resolve the small Manning gradients to ~1e-5.
    assert _fd_check(case, A.misfit_and_grads) < 1e-4


def test_a_planted_friction_mistake_fails(case, monkeypatch) -> None:  # noqa: ANN001
    real_stage = A.stage

    def wrong(m, h, hu, hv, n, dzl, dzr, dt, g, dry, thin=0.0):  # noqa: ANN001, ANN202
        # Manning gradient of the wrong power: n enters as n^2 * n / n0 with n0 detached.
        n_bad = n * n.detach() / n.detach().clamp(min=1e-30) ** 0.5 * n.detach() ** 0.5
        return real_stage(m, h, hu, hv, n + 0.5 * (n_bad - n_bad.detach()), dzl, dzr, dt, g, dry, thin)

    def grads(*a):  # noqa: ANN002, ANN202
        monkeypatch.setattr(A, "stage", wrong)
        try:
            return A.misfit_and_grads(*a)
        finally:
            monkeypatch.setattr(A, "stage", real_stage)

    assert _fd_check(case, grads) > 1e-3
