"""FP64 NumPy restatement of the resident Heun step (TEST INFRASTRUCTURE ONLY).

The same scheme as the reference's GPU-resident kernels, evaluated in float64:

- per-edge words: Audusse face states from the precomputed bed steps
  (``flux.py:146-169``), HLLC (``flux.py:98-140``) with Toro's dry-bed speeds
  (``flux.py:75-95``), the hydrostatic correction (``resident.py:268-271``),
  walls pass no mass (``resident.py:148-151``); an edge dry on both sides
  gives nothing (``resident.py:140-145``);
- per-cell sums in CSR slot order (``resident.py:312-344``);
- stage update: positivity clamp, dry momentum zeroing, film drag, semi-implicit
  Manning (``resident.py:459-483``, ``flux.py:46-51``), Heun average with dry
  cells carrying no momentum (``resident.py:489-501``);
- constant sources after the corrector (``resident.py:695-706``) and their dt
  bound (``resident.py:709-728``);
- the prologue CFL bound from ``wave_bound`` (``resident.py:528-610``).

Compensated depth is the float32 emulation of extra precision, so in float64
it is plain addition. With no sources and a fixed dt this is the scheme of the
reference's own FP64 test oracle ``tests/unit/_cpu_step.py:36-285``, which
sums in edge order and reconstructs from the bed instead of the bed steps;
the two agree to float64 rounding (checked against fixtures made from it).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import NDArray

F = np.float64
FILM_DEPTHS = 4.0


@dataclass
class Mesh64:
    """The mesh in the device edge order, as float64/int64 arrays."""

    cl: NDArray[np.int64]
    cr: NDArray[np.int64]
    nx: NDArray[np.float64]
    ny: NDArray[np.float64]
    length: NDArray[np.float64]
    dzl: NDArray[np.float64]
    dzr: NDArray[np.float64]
    area: NDArray[np.float64]
    cptr: NDArray[np.int64]
    cidx: NDArray[np.int64]

    @property
    def N(self) -> int:  # noqa: N802
        return int(self.area.size)

    @classmethod
    def from_device_arrays(cls, m: dict[str, Any]) -> Mesh64:
        dz = np.asarray(m["dz2"], dtype=F).reshape(-1, 2)
        return cls(
            cl=np.asarray(m["cell_l"], dtype=np.int64),
            cr=np.asarray(m["cell_r"], dtype=np.int64),
            nx=np.asarray(m["nx"], dtype=F),
            ny=np.asarray(m["ny"], dtype=F),
            length=np.asarray(m["len"], dtype=F),
            dzl=dz[:, 0].copy(),
            dzr=dz[:, 1].copy(),
            area=np.asarray(m["area"], dtype=F),
            cptr=np.asarray(m["cptr"], dtype=np.int64),
            cidx=np.asarray(m["cidx"], dtype=np.int64),
        )


@dataclass
class State64:
    h: NDArray[np.float64]
    hu: NDArray[np.float64]
    hv: NDArray[np.float64]

    def copy(self) -> State64:
        return State64(self.h.copy(), self.hu.copy(), self.hv.copy())


def safe_vel(h: NDArray, hq: NDArray, dry: float) -> NDArray:
    """Kurganov-Petrova desingularised velocity (flux.py:63-66)."""
    e = FILM_DEPTHS * dry
    return np.where(h > dry, 2.0 * h * hq / (h * h + np.maximum(h * h, e * e)), 0.0)


def hllc_speeds(hL: NDArray, unL: NDArray, hR: NDArray, unR: NDArray, g: float, dry: float) -> tuple[NDArray, NDArray]:
    """Roe speeds blended with Toro's dry-bed speeds over 1-2 dry_tol (flux.py:75-95)."""
    cL = np.sqrt(g * np.maximum(hL, 0.0))
    cR = np.sqrt(g * np.maximum(hR, 0.0))
    sqL = np.sqrt(np.maximum(hL, 0.0))
    sqR = np.sqrt(np.maximum(hR, 0.0))
    den = sqL + sqR
    ok = den > 1e-12
    u_roe = np.where(ok, (sqL * unL + sqR * unR) / np.where(ok, den, 1.0), 0.0)
    c_roe = np.sqrt(g * 0.5 * (hL + hR))
    sl = np.minimum(unL - cL, u_roe - c_roe)
    sr = np.maximum(unR + cR, u_roe + c_roe)
    wl = np.clip(hL / dry - 1.0, 0.0, 1.0)
    wr = np.clip(hR / dry - 1.0, 0.0, 1.0)
    dl = (1.0 - wl) * wr
    dr = (1.0 - wr) * wl
    roe = 1.0 - dl - dr
    return (
        roe * sl + dl * (unR - 2.0 * cR) + dr * (unL - cL),
        roe * sr + dl * (unR + cR) + dr * (unL + 2.0 * cL),
    )


def hllc(hL, hqnL, hqtL, hR, hqnR, hqtR, g: float, dry: float):  # noqa: ANN001, ANN201
    """HLLC flux in the face frame and the largest wave speed (flux.py:98-140)."""
    unL, utL = safe_vel(hL, hqnL, dry), safe_vel(hL, hqtL, dry)
    unR, utR = safe_vel(hR, hqnR, dry), safe_vel(hR, hqtR, dry)
    SL, SR = hllc_speeds(hL, unL, hR, unR, g, dry)
    num = SL * hR * (unR - SR) - SR * hL * (unL - SL)
    den = hR * (unR - SR) - hL * (unL - SL)
    ok = np.abs(den) > 1e-12
    SM = np.where(ok, num / np.where(ok, den, 1.0), 0.0)
    FL_h, FR_h = hL * unL, hR * unR
    FL_hn = hL * unL * unL + 0.5 * g * hL * hL
    FR_hn = hR * unR * unR + 0.5 * g * hR * hR
    FL_ht, FR_ht = hL * unL * utL, hR * unR * utR
    dSL, dSR = SL - SM, SR - SM
    kL = hL * (SL - unL) / np.where(np.abs(dSL) > 1e-14, dSL, 1e-14)
    kR = hR * (SR - unR) / np.where(np.abs(dSR) > 1e-14, dSR, 1e-14)
    FsL = (FL_h + SL * (kL - hL), FL_hn + SL * (kL * SM - hL * unL), FL_ht + SL * (kL * utL - hL * utL))
    FsR = (FR_h + SR * (kR - hR), FR_hn + SR * (kR * SM - hR * unR), FR_ht + SR * (kR * utR - hR * utR))
    FL = (FL_h, FL_hn, FL_ht)
    FR = (FR_h, FR_hn, FR_ht)
    out = tuple(
        np.where(SL >= 0, a, np.where(SM >= 0, b, np.where(SR >= 0, c, d)))
        for a, b, c, d in zip(FL, FsL, FsR, FR)
    )
    return out[0], out[1], out[2], np.maximum(np.abs(SL), np.abs(SR))


def _sides(m: Mesh64, s: State64) -> tuple[NDArray, ...]:
    """Left/right states of every edge, the wall ghost mirrored (resident.py:126-134)."""
    wall = m.cr < 0
    cri = np.where(wall, m.cl, m.cr)
    hL, huL, hvL = s.h[m.cl], s.hu[m.cl], s.hv[m.cl]
    qn = huL * m.nx + hvL * m.ny
    hR = np.where(wall, hL, s.h[cri])
    huR = np.where(wall, huL - 2.0 * qn * m.nx, s.hu[cri])
    hvR = np.where(wall, hvL - 2.0 * qn * m.ny, s.hv[cri])
    return hL, huL, hvL, hR, huR, hvR, wall


def face_states(m: Mesh64, hL, huL, hvL, hR, huR, hvR, dry: float):  # noqa: ANN001, ANN201
    """Audusse reconstruction from the bed steps, rotated to the face (flux.py:146-169)."""
    hLs = np.maximum(0.0, hL - m.dzl)
    hRs = np.maximum(0.0, hR - m.dzr)
    uL, vL = safe_vel(hL, huL, dry), safe_vel(hL, hvL, dry)
    uR, vR = safe_vel(hR, huR, dry), safe_vel(hR, hvR, dry)
    huLs, hvLs, huRs, hvRs = hLs * uL, hLs * vL, hRs * uR, hRs * vR
    return (
        hLs,
        huLs * m.nx + hvLs * m.ny,
        -huLs * m.ny + hvLs * m.nx,
        hRs,
        huRs * m.nx + hvRs * m.ny,
        -huRs * m.ny + hvRs * m.nx,
    )


def edge_words(m: Mesh64, s: State64, g: float, dry: float) -> NDArray[np.float64]:
    """(6, E): mass, left momentum x/y, right momentum x/y (times length), Smax*len."""
    hL, huL, hvL, hR, huR, hvR, wall = _sides(m, s)
    live = ~((hL < dry) & (hR < dry))
    hLs, hqnL, hqtL, hRs, hqnR, hqtR = face_states(m, hL, huL, hvL, hR, huR, hvR, dry)
    Fh, Fhn, Fht, smax = hllc(hLs, hqnL, hqtL, hRs, hqnR, hqtR, g, dry)
    Fh = np.where(wall, 0.0, Fh)
    Fhu = Fhn * m.nx - Fht * m.ny
    Fhv = Fhn * m.ny + Fht * m.nx
    slg = 0.5 * g * (hL * hL - hLs * hLs)
    srg = 0.5 * g * (hR * hR - hRs * hRs)
    L = m.length
    w = np.stack(
        [
            Fh * L,
            Fhu * L + slg * m.nx * L,
            Fhv * L + slg * m.ny * L,
            Fhu * L + srg * m.nx * L,
            Fhv * L + srg * m.ny * L,
            smax * L,
        ]
    )
    return np.where(live[None, :], w, 0.0)


def cell_sums(m: Mesh64, w: NDArray[np.float64]) -> tuple[NDArray, NDArray, NDArray, NDArray]:
    """Per-cell sums of the edge words in CSR slot order (resident.py:312-344)."""
    n = m.N
    deg = np.diff(m.cptr)
    sdh = np.zeros(n)
    sdhu = np.zeros(n)
    sdhv = np.zeros(n)
    lam = np.zeros(n)
    for p in range(int(deg.max(initial=0))):
        cells = np.flatnonzero(deg > p)
        sl = m.cidx[m.cptr[cells] + p]
        e = sl >> 1
        right = (sl & 1) == 1
        sdh[cells] += np.where(right, -w[0, e], w[0, e])
        sdhu[cells] += np.where(right, -w[3, e], w[1, e])
        sdhv[cells] += np.where(right, -w[4, e], w[2, e])
        lam[cells] += w[5, e]
    return sdh, sdhu, sdhv, lam


def stage(m: Mesh64, s: State64, n: NDArray[np.float64], dt: float, g: float, dry: float) -> tuple[State64, NDArray]:
    """One forward-Euler stage and the per-cell wave-speed sum (resident.py:425-483)."""
    w = edge_words(m, s, g, dry)
    sdh, sdhu, sdhv, lam = cell_sums(m, w)
    dti = dt / m.area
    h = np.maximum(s.h - dti * sdh, 0.0)
    hu = s.hu - dti * sdhu
    hv = s.hv - dti * sdhv
    wet = h >= dry
    hu = np.where(wet, hu, 0.0)
    hv = np.where(wet, hv, 0.0)
    e = FILM_DEPTHS * dry
    hf = np.where(wet & (h < e), h, e)
    drag = 1.0 / (1.0 + dt * ((e * e - hf * hf) / (2.0 * hf * hf) * np.sqrt(g / e)))
    hu = hu * drag
    hv = hv * drag
    hs = np.where(wet, h, 1.0)
    spd = np.sqrt((hu / hs) ** 2 + (hv / hs) ** 2)
    den = np.where(wet, 1.0 + dt * (g * n * n * spd / hs ** (4.0 / 3.0)), 1.0)
    return State64(h, hu / den, hv / den), lam


def source_bound(m: Mesh64, s: State64, src: NDArray[np.float64], g: float, dry: float, cfl: float) -> float:
    """Smallest post-source dt bound of the wet source cells (resident.py:709-728)."""
    cells = np.flatnonzero((src != 0.0) & (s.h >= dry))
    best = np.inf
    for i in cells.tolist():
        u = np.sqrt(s.hu[i] ** 2 + s.hv[i] ** 2) / s.h[i]
        c = np.sqrt(g * s.h[i])
        lam = 0.0
        for k in range(int(m.cptr[i]), int(m.cptr[i + 1])):
            sl = int(m.cidx[k])
            e = sl >> 1
            nb = int(m.cl[e]) if (sl & 1) else int(m.cr[e])
            dry_side = nb >= 0 and s.h[nb] < dry
            lam += (u + (2.0 if dry_side else 1.0) * c) * m.length[e]
        if lam > 1e-30 and m.area[i] > 1e-10:
            best = min(best, cfl * m.area[i] / lam)
    return float(best)


def heun_step(
    m: Mesh64,
    s: State64,
    n: NDArray[np.float64],
    dt: float,
    g: float,
    dry: float,
    cfl: float = 0.8,
    src: NDArray[np.float64] | None = None,
) -> tuple[State64, float]:
    """One Heun step plus sources; returns the state and the next-step CFL bound."""
    s1, _ = stage(m, s, n, dt, g, dry)
    s2, lam = stage(m, s1, n, dt, g, dry)
    h = 0.5 * (s.h + s2.h)
    wet = h >= dry
    hu = np.where(wet, 0.5 * (s.hu + s2.hu), 0.0)
    hv = np.where(wet, 0.5 * (s.hv + s2.hv), 0.0)
    ok = (lam > 1e-30) & (m.area > 1e-10)
    bound = float(np.min(cfl * m.area[ok] / lam[ok])) if ok.any() else np.inf
    out = State64(h, hu, hv)
    if src is not None and np.any(src != 0.0):
        out.h = np.where(src != 0.0, np.maximum(out.h + dt * src, 0.0), out.h)
        bound = min(bound, source_bound(m, out, src, g, dry, cfl))
    return out, bound


def prologue_bound(m: Mesh64, s: State64, g: float, dry: float, cfl: float) -> float:
    """CFL bound the prologue plans the first step from (resident.py:528-610)."""
    hL, huL, hvL, hR, huR, hvR, _ = _sides(m, s)
    hLs, hqnL, _, hRs, hqnR, _ = face_states(m, hL, huL, hvL, hR, huR, hvR, dry)
    SL, SR = hllc_speeds(hLs, safe_vel(hLs, hqnL, dry), hRs, safe_vel(hRs, hqnR, dry), g, dry)
    contrib = np.where((hL >= dry) | (hR >= dry), np.maximum(np.abs(SL), np.abs(SR)) * m.length, 0.0)
    w = np.zeros((6, contrib.size))
    w[5] = contrib
    _, _, _, lam = cell_sums(m, w)
    ok = (lam > 1e-30) & (m.area > 1e-10)
    return float(np.min(cfl * m.area[ok] / lam[ok])) if ok.any() else np.inf


def run_dt_sequence(
    m: Mesh64,
    s0: State64,
    n: NDArray[np.float64],
    dts: list[float] | NDArray[np.float64],
    g: float,
    dry: float,
    src: NDArray[np.float64] | None = None,
) -> State64:
    """Replay a recorded dt sequence (e.g. the FP32 run's) in float64."""
    s = s0.copy()
    for dt in dts:
        s, _ = heun_step(m, s, n, float(dt), g, dry, src=src)
    return s


__all__ = [
    "Mesh64",
    "State64",
    "cell_sums",
    "edge_words",
    "hllc",
    "hllc_speeds",
    "heun_step",
    "prologue_bound",
    "run_dt_sequence",
    "safe_vel",
    "source_bound",
    "stage",
]
