"""Host-driven comparison backends on the HIP kernels, for A/B checks.

The reference keeps two host-driven backends beside its resident one: the
base solver (``solver_impl="baseline"``, ``inundation/vulkan/solver.py:499-760``)
and ``fixed_dt_batch_barrier`` (``vulkan/fixed_dt_batch_barrier.py:68-270``).
Both estimate the CFL dt on the host every ``cfl_interval`` steps, take
``min(0.8 * dt_cfl, dt_max)`` and keep it fixed until the next estimate;
both accumulate fluxes with atomics into a plain float32 depth, with no
active set.

``SWESolverHipHostDriven`` runs those loops over the same stage kernels as
``hip_resident``, with the same defaults (``deterministic = False``,
``compensated = False``, ``active_set = False``; each can be switched back
on). Each run of fixed-dt steps goes to the device clock as a window whose
every step takes the fixed dt: ``dt_max_ticks`` and ``dry_dt_ticks`` are set
to it and the CFL factor is made large enough that the cap always applies.
Simulated time is therefore a count of 2^-20 s ticks and the fixed dt is
rounded down to a whole tick, where the reference keeps float64 seconds.

- ``mode="baseline"``: steps past an output time and snapshots there, once
  per step; ``source_fn`` (a rate per cell as a function of time) is applied
  step by step; momentum snapshots and open boundaries are supported.
- ``mode="fixed_dt_batch_barrier"``: batches of up to ``cfl_interval``
  steps, each step shortened to land on output times (or a snapshot every
  ``output_every_steps`` steps across the solver's lifetime); it refuses
  momentum snapshots and open boundaries, as the reference does.
"""

from __future__ import annotations

import math
import sys
import time
from typing import TYPE_CHECKING, Any

import numpy as np

from ..tuning import (
    CFL_EPSILON_MIN,
    CFL_HOST_DRIVEN,
    CFL_SANITY_MAX,
    dt_to_ticks,
    seconds_to_ticks,
    ticks_to_seconds,
)
from . import lib as L
from .solver import SWESolverHipResident

if TYPE_CHECKING:
    from collections.abc import Callable

    from numpy.typing import NDArray

HOST_CFL_SAFETY = 0.8  # fraction of the CFL dt the host loops take (solver.py:621)
_ALWAYS_CAP = 3.0e38  # CFL factor that makes every step take the fixed dt
MODES = ("baseline", "fixed_dt_batch_barrier")


class SWESolverHipHostDriven(SWESolverHipResident):
    """The reference's host-driven run loops over the HIP stage kernels."""

    default_cfl: float = CFL_HOST_DRIVEN
    deterministic: bool = False
    compensated: bool = False
    active_set: bool = False

    def __init__(self, *args: Any, mode: str = "baseline", **kwargs: Any) -> None:
        if mode not in MODES:
            raise ValueError(f"mode must be one of {MODES}, got {mode!r}")
        self.mode = mode
        if any(np.asarray(e).size for e in kwargs.get("open_edges", ())):
            self.deterministic = True  # open edges here need the fixed-order gather
        super().__init__(*args, **kwargs)

    # -- device primitives ----------------------------------------------------

    def _ctrl(self, now: int, stop: int, dt_ticks: int) -> L.InuCtrl:
        c = L.InuCtrl()
        c.now_ticks = now
        c.stop_ticks = stop
        c.steps = 0
        c.dt_ticks = 0
        c.dt_max_ticks = dt_ticks
        c.dry_dt_ticks = dt_ticks
        c.cfl_safety = _ALWAYS_CAP
        c.last_raw = 1.0  # any bound in range: the cap decides
        return c

    def _cfl_estimate(self, now: int) -> float:
        """The CFL dt of the current state (the prologue's bound; 1e30 on an all-dry mesh)."""
        self._ctx.ctrl_write(self._ctrl(now, now + 1, 1))
        self._ctx.prologue()
        return float(self._ctx.ctrl_read().last_raw)

    def _advance(self, now: int, dt_ticks: int, n: int, stop: int) -> tuple[int, int]:
        """n steps of dt_ticks from now, the last one shortened to stop: (new time, steps taken)."""
        self._ctx.ctrl_write(self._ctrl(now, stop, dt_ticks))
        self._ctx.retarget()
        # Batches of one length repeat, so a recorded graph pays off from 16 steps.
        self._ctx.steps(n, use_graph=n >= 16)
        c = self._ctx.ctrl_read()
        if int(c.now_ticks) == now:
            raise RuntimeError(
                f"host-driven run made no time progress at t={ticks_to_seconds(now):.6f}s "
                "(the next CFL bound fell below 1e-10 s)"
            )
        return int(c.now_ticks), int(c.steps)

    # -- run --------------------------------------------------------------------

    def run(
        self,
        t_end: float,
        output_interval_s: float,
        dt_max: float,
        dt_init: float,
        cfl_interval: int,
        progress: bool = True,
        source_rate: NDArray[np.float32] | None = None,
        source_fn: Callable[[float], NDArray[np.float32]] | None = None,
        resume: bool = False,
        t_start: float = 0.0,
        capture_momentum_snapshots: bool = False,
        output_every_steps: int | None = None,
    ) -> Any:
        """Advance to ``t_end`` with host-chosen fixed dt; same arguments and returns as ``hip_resident``."""
        batch = self.mode == "fixed_dt_batch_barrier"
        if batch and capture_momentum_snapshots:
            raise ValueError("fixed_dt_batch_barrier does not capture momentum snapshots")
        if batch and self.n_open:
            raise ValueError("fixed_dt_batch_barrier does not record boundary volumes")
        if batch and source_fn is not None:
            raise RuntimeError("fixed_dt_batch_barrier requires GPU source kernel for batched source terms")
        if cfl_interval < 1:
            raise ValueError(f"cfl_interval must be positive, got {cfl_interval}")
        self._apply_params()
        if not resume:
            self._reset_to_initial()
            self.open_volume_snapshots = []
        if self.compensated or self.n_open:
            self._check_error_free_arithmetic()
        self._set_sources(None if source_fn is not None else source_rate)

        start = seconds_to_ticks(t_start if resume else 0.0)
        end = seconds_to_ticks(t_end)
        interval = seconds_to_ticks(output_interval_s)
        snaps = [self.download_h() if resume else self._h0.copy()]
        times = [ticks_to_seconds(start)]
        hu_s: list[NDArray[np.float32]] = []
        hv_s: list[NDArray[np.float32]] = []

        def capture(t: int) -> bool:
            h = self.download_h()
            finite = bool(np.isfinite(h).all())
            snaps.append(h if finite else np.nan_to_num(h, nan=0.0, posinf=0.0))
            times.append(ticks_to_seconds(t))
            if capture_momentum_snapshots:
                hu_s.append(self.download_hu())
                hv_s.append(self.download_hv())
            if self.n_open:
                self.open_volume_snapshots.append(self.download_open_volumes())
            return finite

        if capture_momentum_snapshots:
            hu_s.append(self.download_hu())
            hv_s.append(self.download_hv())
        if self.n_open:
            self.open_volume_snapshots.append(self.download_open_volumes())

        t = start
        step = 0
        last_log = time.monotonic()
        dt = float(dt_init)
        next_out = start + interval
        finite = True
        while t < end:
            dt = min(dt, dt_max)
            if step % cfl_interval == 0:
                dt_cfl = self._cfl_estimate(t)
                if dt_cfl < CFL_EPSILON_MIN:
                    if progress:
                        sys.stdout.write(f"[solver][step {step}] CFL dt too small ({dt_cfl:.3e}), stopping\n")
                    break
                if dt_cfl < CFL_SANITY_MAX:
                    dt = min(dt_cfl * HOST_CFL_SAFETY, dt_max)
            dtt = dt_to_ticks(dt)
            n = cfl_interval - step % cfl_interval
            if source_fn is not None:
                n = 1
                rate = np.asarray(source_fn(ticks_to_seconds(t)), dtype=np.float32)
                self._set_sources(rate)
            stop = end
            if batch and output_every_steps is not None:
                n = min(n, output_every_steps - (self.steps_total + step) % output_every_steps)
            elif batch:
                stop = min(end, next_out)  # steps land on the output time
            n = min(n, math.ceil((stop - t) / dtt))
            if not batch:
                # baseline: snapshot at the first step at or past the output time
                n = min(n, max(1, math.ceil((next_out - t) / dtt)))
            t, taken = self._advance(t, dtt, n, min(stop, t + n * dtt))
            step += taken
            if batch and output_every_steps is not None:
                due = (self.steps_total + step) % output_every_steps == 0 or t >= end
            else:
                due = t >= next_out
            if due:
                finite = capture(t)
                if not finite:
                    break
                if not (batch and output_every_steps is not None):
                    next_out += interval
            now_wall = time.monotonic()
            if progress and now_wall - last_log >= 10.0:
                sys.stdout.write(f"[solver][host {step}] sim={ticks_to_seconds(t):.2f}/{t_end:.2f}s dt={dt:.4f}\n")
                sys.stdout.flush()
                last_log = now_wall
        self.steps_total += step
        if finite:
            h = self.download_h()
            if not np.allclose(snaps[-1], h):
                capture(t)
        if capture_momentum_snapshots:
            return snaps, times, hu_s, hv_s
        return snaps, times


__all__ = ["MODES", "SWESolverHipHostDriven"]
