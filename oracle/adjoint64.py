"""FP64 reference gradients of a gauge-depth misfit (TEST INFRASTRUCTURE ONLY).

The resident Heun scheme of ``oracle/ref64.py`` restated in torch float64 on
the CPU, with the bed elevation z and the per-cell Manning n as inputs, so that
reverse-mode automatic differentiation gives the exact FP64 gradient of

    loss = sum over samples k and gauges i of (h_i(t_k) - obs_ik)^2

for a replayed dt sequence. It is independent of the hand-written HIP adjoint:
it differentiates the FP64 scheme mechanically. The same choices as the HIP
adjoint at the scheme's kinks:

- bed steps from z (reference ``mesh/geometry.py:241-254``): the face sits on
  the higher bed, the left one on a tie; walls have no step;
- the square root of a non-positive depth has derivative 0 (a dry side);
- wet/dry tests, clamps and the HLLC branch choice take the branch the
  forward values take;
- held (no derivative through them): the weights of the wet/dry speed
  blend, the star state of a degenerate HLLC star region (|S - SM| at the
  1e-14 guard), and, below ``thin`` face depth, a side's velocity and
  celerity (its depth still enters).

Scheme (as ``oracle/ref64.py``): Audusse face depths ``max(0, h - dz)``,
desingularised velocities (reference ``flux.py:63-66``), HLLC with blended
dry-bed speeds (``flux.py:75-140``), hydrostatic correction, dry-dry edges
give nothing, CSR-order sums, clamp, dry momentum zeroing, film drag and
semi-implicit Manning (``resident.py:459-483``), Heun average, constant
sources after the corrector (``resident.py:695-706``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

F64 = torch.float64
FILM_DEPTHS = 4.0


@dataclass
class MeshT:
    cl: torch.Tensor
    cr: torch.Tensor  # -1 for walls
    crs: torch.Tensor  # cr with walls replaced by cl
    wall: torch.Tensor
    nx: torch.Tensor
    ny: torch.Tensor
    length: torch.Tensor
    area: torch.Tensor
    slot_cell: torch.Tensor  # CSR slot -> cell
    slot_edge: torch.Tensor  # CSR slot -> edge
    slot_right: torch.Tensor  # CSR slot is the edge's right side

    @property
    def N(self) -> int:  # noqa: N802
        return int(self.area.numel())

    @classmethod
    def from_device_arrays(cls, m: dict[str, Any]) -> MeshT:
        cl = torch.as_tensor(np.asarray(m["cell_l"], np.int64))
        cr = torch.as_tensor(np.asarray(m["cell_r"], np.int64))
        wall = cr < 0
        cptr = np.asarray(m["cptr"], np.int64)
        cidx = np.asarray(m["cidx"], np.int64)
        return cls(
            cl=cl,
            cr=cr,
            crs=torch.where(wall, cl, cr),
            wall=wall,
            nx=torch.as_tensor(np.asarray(m["nx"], np.float64)),
            ny=torch.as_tensor(np.asarray(m["ny"], np.float64)),
            length=torch.as_tensor(np.asarray(m["len"], np.float64)),
            area=torch.as_tensor(np.asarray(m["area"], np.float64)),
            slot_cell=torch.as_tensor(np.repeat(np.arange(cptr.size - 1), np.diff(cptr))),
            slot_edge=torch.as_tensor(cidx >> 1),
            slot_right=torch.as_tensor((cidx & 1) == 1),
        )


def bed_steps(m: MeshT, z: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    zl = z[m.cl]
    zr = torch.where(m.wall, zl, z[m.crs])
    left_high = zl >= zr
    dzl = torch.where(left_high, torch.zeros_like(zl), zr - zl)
    dzr = torch.where(left_high, zl - zr, torch.zeros_like(zl))
    return dzl, dzr


def _sqrt0(x: torch.Tensor) -> torch.Tensor:
    pos = x > 0.0
    return torch.where(pos, torch.sqrt(torch.where(pos, x, torch.ones_like(x))), torch.zeros_like(x))


def _vel(h: torch.Tensor, q: torch.Tensor, dry: float) -> torch.Tensor:
    e = FILM_DEPTHS * dry
    hh = h * h
    return torch.where(h > dry, 2.0 * h * q / (hh + torch.clamp(hh, min=e * e)), torch.zeros_like(h))


def _hold(x: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
    """x with no derivative where ``where`` holds."""
    return torch.where(where, x.detach(), x)


def _speeds(hL, unL, hR, unR, g: float, dry: float, thinL=None, thinR=None):  # noqa: ANN001, ANN202
    sqL, sqR = _sqrt0(hL), _sqrt0(hR)
    hLc, hRc = hL, hR
    if thinL is not None:
        sqL, sqR = _hold(sqL, thinL), _hold(sqR, thinR)
        hLc, hRc = _hold(hL, thinL), _hold(hR, thinR)
    cL, cR = np.sqrt(g) * sqL, np.sqrt(g) * sqR
    den = sqL + sqR
    ok = den > 1e-12
    u_roe = torch.where(ok, (sqL * unL + sqR * unR) / torch.where(ok, den, torch.ones_like(den)), torch.zeros_like(den))
    c_roe = _sqrt0(0.5 * g * (hLc + hRc))
    sl = torch.minimum(unL - cL, u_roe - c_roe)
    sr = torch.maximum(unR + cR, u_roe + c_roe)
    # The blend weights are the wet/dry decision: held.
    wl = torch.clamp(hL / dry - 1.0, 0.0, 1.0).detach()
    wr = torch.clamp(hR / dry - 1.0, 0.0, 1.0).detach()
    dl = (1.0 - wl) * wr
    dr = (1.0 - wr) * wl
    roe = 1.0 - dl - dr
    return (roe * sl + dl * (unR - 2.0 * cR) + dr * (unL - cL), roe * sr + dl * (unR + cR) + dr * (unL + 2.0 * cL))


def _hllc(hL, hqnL, hqtL, hR, hqnR, hqtR, g: float, dry: float, thin: float = 0.0):  # noqa: ANN001, ANN202
    unL, utL = _vel(hL, hqnL, dry), _vel(hL, hqtL, dry)
    unR, utR = _vel(hR, hqnR, dry), _vel(hR, hqtR, dry)
    thinL = thinR = None
    if thin > 0.0:
        thinL, thinR = hL < thin, hR < thin
        unL, utL, unR, utR = _hold(unL, thinL), _hold(utL, thinL), _hold(unR, thinR), _hold(utR, thinR)
    SL, SR = _speeds(hL, unL, hR, unR, g, dry, thinL, thinR)
    num = SL * hR * (unR - SR) - SR * hL * (unL - SL)
    den = hR * (unR - SR) - hL * (unL - SL)
    ok = den.abs() > 1e-12
    SM = torch.where(ok, num / torch.where(ok, den, torch.ones_like(den)), torch.zeros_like(den))
    FL = (hL * unL, hL * unL * unL + 0.5 * g * hL * hL, hL * unL * utL)
    FR = (hR * unR, hR * unR * unR + 0.5 * g * hR * hR, hR * unR * utR)
    dSL, dSR = SL - SM, SR - SM
    okL, okR = dSL.abs() > 1e-14, dSR.abs() > 1e-14
    kL = _hold(hL * (SL - unL) / torch.where(okL, dSL, torch.full_like(dSL, 1e-14)), ~okL)
    kR = _hold(hR * (SR - unR) / torch.where(okR, dSR, torch.full_like(dSR, 1e-14)), ~okR)
    FsL = (FL[0] + SL * (kL - hL), FL[1] + SL * (kL * SM - hL * unL), FL[2] + SL * (kL * utL - hL * utL))
    FsR = (FR[0] + SR * (kR - hR), FR[1] + SR * (kR * SM - hR * unR), FR[2] + SR * (kR * utR - hR * utR))
    out = [
        torch.where(SL >= 0, a, torch.where(SM >= 0, b, torch.where(SR >= 0, c, d)))
        for a, b, c, d in zip(FL, FsL, FsR, FR)
    ]
    return out[0], out[1], out[2]


def edge_words(m: MeshT, h, hu, hv, dzl, dzr, g: float, dry: float, thin: float = 0.0):  # noqa: ANN001, ANN201
    """(5, E) words: mass, left momentum x/y, right momentum x/y (times length)."""
    hL, huL, hvL = h[m.cl], hu[m.cl], hv[m.cl]
    qn = huL * m.nx + hvL * m.ny
    hR = torch.where(m.wall, hL, h[m.crs])
    huR = torch.where(m.wall, huL - 2.0 * qn * m.nx, hu[m.crs])
    hvR = torch.where(m.wall, hvL - 2.0 * qn * m.ny, hv[m.crs])
    live = ~((hL < dry) & (hR < dry))
    hLs = torch.clamp(hL - dzl, min=0.0)
    hRs = torch.clamp(hR - dzr, min=0.0)
    uL, vL = _vel(hL, huL, dry), _vel(hL, hvL, dry)
    uR, vR = _vel(hR, huR, dry), _vel(hR, hvR, dry)
    huLs, hvLs, huRs, hvRs = hLs * uL, hLs * vL, hRs * uR, hRs * vR
    Fh, Fhn, Fht = _hllc(hLs, huLs * m.nx + hvLs * m.ny, -huLs * m.ny + hvLs * m.nx,
                         hRs, huRs * m.nx + hvRs * m.ny, -huRs * m.ny + hvRs * m.nx, g, dry, thin)
    Fh = torch.where(m.wall, torch.zeros_like(Fh), Fh)
    Fhu = Fhn * m.nx - Fht * m.ny
    Fhv = Fhn * m.ny + Fht * m.nx
    slg = 0.5 * g * (hL * hL - hLs * hLs)
    srg = 0.5 * g * (hR * hR - hRs * hRs)
    L = m.length
    w = torch.stack([Fh * L, (Fhu + slg * m.nx) * L, (Fhv + slg * m.ny) * L, (Fhu + srg * m.nx) * L,
                     (Fhv + srg * m.ny) * L])
    return torch.where(live[None, :], w, torch.zeros_like(w))


def stage(m: MeshT, h, hu, hv, n, dzl, dzr, dt: float, g: float, dry: float, thin: float = 0.0):  # noqa: ANN001, ANN201
    w = edge_words(m, h, hu, hv, dzl, dzr, g, dry, thin)
    e = m.slot_edge
    r = m.slot_right
    sdh = torch.zeros_like(h).index_add(0, m.slot_cell, torch.where(r, -w[0, e], w[0, e]))
    sdhu = torch.zeros_like(h).index_add(0, m.slot_cell, torch.where(r, -w[3, e], w[1, e]))
    sdhv = torch.zeros_like(h).index_add(0, m.slot_cell, torch.where(r, -w[4, e], w[2, e]))
    dti = dt / m.area
    h1 = torch.clamp(h - dti * sdh, min=0.0)
    hu1 = hu - dti * sdhu
    hv1 = hv - dti * sdhv
    wet = h1 >= dry
    zero = torch.zeros_like(h1)
    hu1 = torch.where(wet, hu1, zero)
    hv1 = torch.where(wet, hv1, zero)
    ef = FILM_DEPTHS * dry
    hf = torch.where(wet & (h1 < ef), h1, torch.full_like(h1, ef))
    drag = 1.0 / (1.0 + dt * ((ef * ef - hf * hf) / (2.0 * hf * hf) * np.sqrt(g / ef)))
    hu1 = hu1 * drag
    hv1 = hv1 * drag
    hs = torch.where(wet, h1, torch.ones_like(h1))
    spd = _sqrt0((hu1 / hs) ** 2 + (hv1 / hs) ** 2)
    den = torch.where(wet, 1.0 + dt * (g * n * n * spd / hs ** (4.0 / 3.0)), torch.ones_like(h1))
    return h1, hu1 / den, hv1 / den


def heun_step(m: MeshT, h, hu, hv, n, dzl, dzr, dt: float, g: float, dry: float, src=None,  # noqa: ANN001, ANN201
              thin: float = 0.0):
    h1, hu1, hv1 = stage(m, h, hu, hv, n, dzl, dzr, dt, g, dry, thin)
    h2, hu2, hv2 = stage(m, h1, hu1, hv1, n, dzl, dzr, dt, g, dry, thin)
    ho = 0.5 * (h + h2)
    wet = ho >= dry
    zero = torch.zeros_like(ho)
    huo = torch.where(wet, 0.5 * (hu + hu2), zero)
    hvo = torch.where(wet, 0.5 * (hv + hv2), zero)
    if src is not None:
        ho = torch.where(src != 0.0, torch.clamp(ho + dt * src, min=0.0), ho)
    return ho, huo, hvo


def misfit_and_grads(
    m: dict[str, Any],
    z: np.ndarray,
    n: np.ndarray,
    state0: tuple[np.ndarray, np.ndarray, np.ndarray],
    dts: np.ndarray,
    src: np.ndarray | None,
    gauges: np.ndarray,
    obs: np.ndarray,
    sample_every: int,
    g: float = 9.81,
    dry: float = 1e-4,
    thin: float = 0.0,
) -> tuple[float, np.ndarray, np.ndarray]:
    """Loss and its FP64 gradients with respect to z and n.

    ``obs`` has one row per sample: after steps sample_every, 2 sample_every, ...
    """
    mesh = MeshT.from_device_arrays(m)
    zt = torch.tensor(np.asarray(z, np.float64), requires_grad=True)
    nt = torch.tensor(np.asarray(n, np.float64), requires_grad=True)
    h, hu, hv = (torch.as_tensor(np.asarray(a, np.float64)) for a in state0)
    srct = None if src is None else torch.as_tensor(np.asarray(src, np.float64))
    gi = torch.as_tensor(np.asarray(gauges, np.int64))
    ot = torch.as_tensor(np.asarray(obs, np.float64))
    dzl, dzr = bed_steps(mesh, zt)
    loss = torch.zeros((), dtype=F64)
    k = 0
    for step, dt in enumerate(np.asarray(dts, np.float64).tolist(), start=1):
        h, hu, hv = heun_step(mesh, h, hu, hv, nt, dzl, dzr, dt, g, dry, srct, thin)
        if step % sample_every == 0:
            loss = loss + ((h[gi] - ot[k]) ** 2).sum()
            k += 1
    loss.backward()
    return float(loss.detach()), zt.grad.numpy().copy(), nt.grad.numpy().copy()


def gauge_series(
    m: dict[str, Any],
    z: np.ndarray,
    n: np.ndarray,
    state0: tuple[np.ndarray, np.ndarray, np.ndarray],
    dts: np.ndarray,
    src: np.ndarray | None,
    gauges: np.ndarray,
    sample_every: int,
    g: float = 9.81,
    dry: float = 1e-4,
) -> np.ndarray:
    """FP64 gauge depths after every sample_every steps (for finite differences)."""
    with torch.no_grad():
        mesh = MeshT.from_device_arrays(m)
        zt = torch.as_tensor(np.asarray(z, np.float64))
        nt = torch.as_tensor(np.asarray(n, np.float64))
        h, hu, hv = (torch.as_tensor(np.asarray(a, np.float64)) for a in state0)
        srct = None if src is None else torch.as_tensor(np.asarray(src, np.float64))
        dzl, dzr = bed_steps(mesh, zt)
        out = []
        for step, dt in enumerate(np.asarray(dts, np.float64).tolist(), start=1):
            h, hu, hv = heun_step(mesh, h, hu, hv, nt, dzl, dzr, dt, g, dry, srct)
            if step % sample_every == 0:
                out.append(h[torch.as_tensor(np.asarray(gauges, np.int64))].numpy().copy())
        return np.array(out)


__all__ = ["MeshT", "bed_steps", "edge_words", "gauge_series", "heun_step", "misfit_and_grads", "stage"]
