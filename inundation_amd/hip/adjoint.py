"""Gradients of sampled gauge depths with respect to Manning n and bed elevation.

``gauge_depths(n, z, run)`` is a ``torch.autograd.Function``: its forward runs
the HIP solver from ``run``'s initial state (CFL-driven steps, recording each
step's dt, gauge depths every ``every`` steps and checkpoints every ``seg``
steps, or only its start under binomial checkpointing, ``slots`` > 0); its
backward runs the hand-written HIP adjoint (``inu_adj_backward``), which
replays each segment from a checkpoint with the recorded dt sequence and
source rates and runs it in reverse, following the forward's wet/dry and
HLLC branches. The reference has no gradient path; the forward is the same
resident Heun step as ``hip_resident`` (reference
``inundation/vulkan/gpu_resident_batch.py``).
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field

import numpy as np
import torch
from numpy.typing import NDArray

from ..tuning import dt_to_ticks
from . import lib as L


@dataclass
class GaugeRun:
    """One differentiable forward run on a one-scenario context that holds the mesh."""

    ctx: L.Context
    state0: tuple[NDArray[np.float32], NDArray[np.float32], NDArray[np.float32]]
    gauges: NDArray[np.uint32]
    every: int
    n_steps: int
    seg: int = 100
    # checkpoints: 0 keeps one every seg steps; s >= 1 uses the binomial
    # schedule with s checkpoint slots (less memory, more replayed steps)
    slots: int = 0
    g: float = 9.81
    dry_tol: float = 1e-4
    cfl: float = 0.8
    dt_max: float = 1.0
    dt_init: float = 0.1
    flags: int = L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET
    source_cells: NDArray[np.uint32] | None = None
    # (first step, rates of source_cells) per phase, the first at step 0
    phases: list[tuple[int, NDArray[np.float32]]] = field(default_factory=list)
    # or rates as a function of simulated time (s), looked up every graph_steps steps
    rates_at: Callable[[float], NDArray[np.float32]] | None = None
    graph_steps: int = 50

    @property
    def n_cells(self) -> int:
        return int(self.state0[0].size)

    def _start(self, n: NDArray[np.float32], z: NDArray[np.float32]) -> NDArray[np.uint32]:
        c = self.ctx
        c.adj_end()  # sources may not be replaced while a recording is open
        c.set_params(self.g, self.dry_tol, self.cfl, self.flags)
        c.set_manning(np.asarray(n, np.float32))
        c.set_bed(np.asarray(z, np.float32))
        c.set_state(*self.state0, None)
        cells = np.empty(0, np.uint32) if self.source_cells is None else self.source_cells
        if self.rates_at is not None:
            rates0 = self.rates_at(0.0)
        else:
            rates0 = self.phases[0][1] if self.phases else np.empty(0, np.float32)
        c.set_sources(cells, rates0)
        ctrl = L.InuCtrl()
        ctrl.now_ticks = 0
        ctrl.stop_ticks = 1 << 60
        ctrl.steps = 0
        ctrl.dt_ticks = 0
        ctrl.dt_max_ticks = dt_to_ticks(self.dt_max)
        ctrl.dry_dt_ticks = dt_to_ticks(min(self.dt_init, self.dt_max))
        ctrl.cfl_safety = 1.0
        c.ctrl_write(ctrl)
        c.prologue()
        return cells

    def _drive(self, cells: NDArray[np.uint32], step: Callable[[int], None]) -> None:
        """Run n_steps in chunks of at most graph_steps, changing the source rates between chunks."""
        c = self.ctx
        if self.rates_at is not None:
            last = self.rates_at(0.0)
            done = 0
            while done < self.n_steps:
                if done:
                    r = self.rates_at(c.ctrl_read().now_ticks / float(1 << 20))
                    if not np.array_equal(r, last):
                        c.set_source_rates(cells, r)
                        last = r
                m = min(self.graph_steps, self.n_steps - done)
                step(m)
                done += m
            return
        starts = [p[0] for p in self.phases[1:]] + [self.n_steps]
        done = 0
        for k, end in enumerate(starts):
            if k > 0:
                c.set_source_rates(cells, self.phases[k][1])
            while done < end:
                m = min(self.graph_steps, end - done)
                step(m)
                done += m

    def forward(self, n: NDArray[np.float32], z: NDArray[np.float32]) -> NDArray[np.float32]:
        """The recorded run: gauge depths (samples x gauges); the adjoint's tape is kept on the device."""
        cells = self._start(n, z)
        self.ctx.adj_set_slots(self.slots)
        self.ctx.adj_begin(self.n_steps, self.seg, self.gauges, self.every)
        self._drive(cells, self.ctx.adj_steps)
        return self.ctx.adj_series()

    def forward_plain(self, n: NDArray[np.float32], z: NDArray[np.float32]) -> None:
        """The same run without recording (the plain resident step): the cost the adjoint is compared with.

        Ends the context's recording.
        """
        cells = self._start(n, z)
        self._drive(cells, self.ctx.steps)

    def backward(self, seeds: NDArray[np.float32]) -> tuple[NDArray[np.float32], NDArray[np.float32]]:
        gn, gz, _ = self.ctx.adj_backward(seeds, self.n_cells)
        return gn, gz

    def replay(self, n: NDArray[np.float32], z: NDArray[np.float32]) -> NDArray[np.float32]:
        """The recorded steps again (same dt and rates) with other n and bed: for finite differences."""
        self.ctx.set_manning(np.asarray(n, np.float32))
        self.ctx.set_bed(np.asarray(z, np.float32))
        return self.ctx.adj_replay()

    def dts(self) -> NDArray[np.float32]:
        return self.ctx.adj_dts(self.n_steps)


class _GaugeDepths(torch.autograd.Function):
    @staticmethod
    def forward(ctx, n: torch.Tensor, z: torch.Tensor, run: GaugeRun) -> torch.Tensor:  # noqa: ANN001
        series = run.forward(n.detach().cpu().numpy(), z.detach().cpu().numpy())
        ctx.run = run
        return torch.from_numpy(series)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):  # noqa: ANN001, ANN205
        gn, gz = ctx.run.backward(grad.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(gn), torch.from_numpy(gz), None


def gauge_depths(n: torch.Tensor, z: torch.Tensor, run: GaugeRun) -> torch.Tensor:
    """Gauge depths (samples x gauges) of the run with Manning n and bed z, differentiable in both."""
    return _GaugeDepths.apply(n, z, run)


__all__ = ["GaugeRun", "gauge_depths"]
