#!/usr/bin/env python3
"""Benchmark of the HIP resident Heun step on one MI355X.

Workload: the 16 003 008-cell rainfall flood (mixed triangle/quad lattice over
8 km of fBm terrain, 50 mm/h uniform rain from a dry start, per-cell Manning
n, FP32 with compensated depth, active set on, 2000 Heun steps). A "step" is
one Heun step of the whole mesh.

- ``value``: cell-steps/s over the timed steps, all cells counted every step
  (the reference's convention, flood_propagation.py:645-646), with the state
  in HBM before the timed region starts;
- ``roofline``: the corrector stage kernel (the larger of the two stage
  kernels), algorithmic bytes of the active blocks (28 B per edge, 4 B per
  CSR slot, 60 B per cell with the compensated low words) over its
  HIP-event-timed duration, against 8 TB/s;
- ``cpu_baseline``: the FP32 oracle (C++/OpenMP) on the same mesh and state
  for a bounded number of steps, checked bit for bit against the GPU on the
  same steps.

Usage: ``python bench.py [--gpus 1] [--steps 2000] [--warmup 20]``.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s & %HBM roofline: plain vs batched vs adjoint on 1 MI355X"
HBM_PEAK_GBS = 8000.0
# HBM bytes per corrector launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (each scaled by the counters' reading for a 1 GiB plain copy) on
# the same workload: profiles/k_stage_rainfall16M_traffic.md. Re-measured
# whenever the stage kernels change.
CORRECTOR_TRAFFIC = {
    "bytes_per_launch": 1.732e9,
    "traffic_over_algorithmic": 0.977,
    "source": "profiles/k_stage_rainfall16M_traffic.md: mean of 300 launches from the dry start, "
    "95.0 % of cell-steps visited",
}


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def start(ctx, wl, t_stop: float, state=None, now_ticks: int = 0) -> None:  # noqa: ANN001
    from inundation_amd.hip.lib import InuCtrl
    from inundation_amd.tuning import dt_to_ticks, seconds_to_ticks

    if state is None:
        ctx.set_state(wl.h0, wl.hu0, wl.hv0, None)
    else:
        ctx.set_state(*state)
    c = InuCtrl()
    c.now_ticks = now_ticks
    c.stop_ticks = now_ticks + seconds_to_ticks(t_stop)
    c.steps = 0
    c.dt_ticks = 0
    c.dt_max_ticks = dt_to_ticks(wl.dt_max)
    c.dry_dt_ticks = dt_to_ticks(min(wl.dt_init, wl.dt_max))
    c.cfl_safety = 1.0
    ctx.ctrl_write(c)
    ctx.prologue()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=int, default=1, help="shrink the lattice by this factor (testing)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument(
        "--copy-calibration",
        type=int,
        default=0,
        metavar="MIB",
        help="after the timed run, launch the library's plain copy kernel (k_stream_copy) over MIB MiB so "
        "that HBM byte counters can be scaled against a kernel of known traffic",
    )
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one MI355X: the solver has no multi-GPU path "
            "(sharding by Hilbert block ranges is described in DESIGN.md); run with --gpus 1",
            file=sys.stderr,
        )
        return 2

    from inundation_amd.hip import lib as L
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workloads import rainfall

    t0 = time.perf_counter()
    wl = rainfall(scale=args.scale)
    geom = wl.geom
    log(f"mesh {geom.N} cells, {geom.E} edges built in {time.perf_counter() - t0:.1f}s")
    m = device_mesh_arrays(geom)
    ctx = L.Context(0)
    t0 = time.perf_counter()
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    info = ctx.mesh_info()
    log(f"device layout in {time.perf_counter() - t0:.1f}s: {info}")
    flags = L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET
    ctx.set_manning(wl.manning)
    ctx.set_params(9.81, wl.dry_tol, wl.cfl, flags)
    cells, rates = wl.source_pairs()
    ctx.set_sources(cells, rates)
    t_stop = 1800.0  # 30 min of rain; dt_max keeps 2000 steps inside it

    # Warm-up: the same run from the dry start, then the timed run from it again.
    start(ctx, wl, t_stop)
    if args.warmup:
        ctx.steps_timed(args.warmup)
    start(ctx, wl, t_stop)
    ctx.activity(reset=True)
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    ms_pred, ms_corr, ms_span = ctx.steps_timed(args.steps)
    wall = time.perf_counter() - t0
    act = ctx.activity()
    ctrl = ctx.ctrl_read()
    if act["steps"] != args.steps:
        raise RuntimeError(f"{args.steps - act['steps']} of the timed steps were no-ops")
    n, e, s = geom.N, geom.E, int(m["cptr"][-1])
    value = args.steps * n / wall
    # Algorithmic bytes (FP32/int32, each array once per stage): per stage
    # 28 B/edge + 4 B/slot + 4 B/cell CSR pointer + 8 B/cell statics + 12 B/cell
    # stage input + 12 B/cell output; the corrector also reads the start state
    # (12 B/cell); compensated low words 8 B/cell (predictor) and 12 B/cell
    # (corrector). Counted over active blocks only; uniform rain adds nothing.
    corr_bytes = 28 * act["edges"] + 4 * act["slots"] + (36 + 12 + 12) * act["cells"]
    step_bytes = 56 * act["edges"] + 8 * act["slots"] + 104 * act["cells"]
    corr_gbs = corr_bytes / (ms_corr * 1e-3) / 1e9
    step_gbs = step_bytes / ((ms_pred + ms_corr) * 1e-3) / 1e9
    active_frac = act["cells"] / (args.steps * n)
    log(
        f"wall {wall:.3f}s ({wall / args.steps * 1e3:.3f} ms/step), events: predictor {ms_pred / args.steps:.3f} ms, "
        f"corrector {ms_corr / args.steps:.3f} ms, span {ms_span / args.steps:.3f} ms; "
        f"active {active_frac:.3f}; corrector {corr_gbs:.0f} GB/s; t_sim {ctrl.now_ticks / 2**20:.1f}s"
    )

    # Attainable HBM rate: a plain copy of 2 GiB (read + write), for context.
    copy_ms = ctx.stream_copy(2 << 30, 5)
    copy_gbs = 2 * (2 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"plain copy of 2 GiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s read+write")
    if args.copy_calibration:
        ms = ctx.stream_copy(args.copy_calibration << 20, 3)
        log(f"copy calibration: 4 launches of k_stream_copy over {args.copy_calibration} MiB, {ms:.3f} ms each")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, wl, m, info, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "cells": n,
            "edges": e,
            "csr_slots": s,
            "seed": wl.seed,
            "compensated": True,
            "active_set": True,
            "dt_max_s": wl.dt_max,
            "cfl": wl.cfl,
            "simulated_s": ctrl.now_ticks / 2**20,
            "active_cell_fraction": active_frac,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_stage<1,true> (corrector)",
            "achieved": corr_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": corr_gbs / HBM_PEAK_GBS,
            "traffic": CORRECTOR_TRAFFIC["bytes_per_launch"],
            "traffic_over_algorithmic": CORRECTOR_TRAFFIC["traffic_over_algorithmic"],
            "traffic_source": CORRECTOR_TRAFFIC["source"],
            "ms_per_launch": ms_corr / args.steps,
            "attainable_copy_GBs": copy_gbs,
            "algorithmic_bytes_per_launch": corr_bytes / args.steps,
        },
        "heun_step_roofline": {
            "achieved": step_gbs,
            "frac": step_gbs / HBM_PEAK_GBS,
            "unit": "GB/s",
            "bytes_per_step": step_bytes / args.steps,
            "stage_kernels_ms_per_step": (ms_pred + ms_corr) / args.steps,
            "active_cell_steps_per_s_over_stage_kernels": act["cells"] / ((ms_pred + ms_corr) * 1e-3),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()
    return 0


def cpu_baseline(ctx, wl, m, info, budget_s: float) -> dict:  # noqa: ANN001
    """The FP32 oracle on the same mesh and state, checked bitwise against the GPU."""
    from inundation_amd.hip import lib as L
    from oracle.cpu_ref import CpuRef, max_threads, set_threads

    n = wl.geom.N
    state = tuple(ctx.download(f, n) for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO))
    ctrl = ctx.ctrl_read()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    set_threads(cores)
    ref = CpuRef(m)
    flags = L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET
    ref.set_params(9.81, wl.dry_tol, wl.cfl, flags)
    ref.set_manning(wl.manning)
    cells, rates = wl.source_pairs()
    ref.set_sources(cells, rates)
    for be in (ref, ctx):
        start(be, wl, 1800.0, state=state, now_ticks=int(ctrl.now_ticks))
    t0 = time.perf_counter()
    ref.steps(1)
    one = time.perf_counter() - t0
    k = max(1, min(200, int(budget_s / max(one, 1e-6)) - 1))
    t0 = time.perf_counter()
    ref.steps(k)
    dt = time.perf_counter() - t0
    ctx.steps(k + 1, use_graph=False)
    same = all(
        np.array_equal(ctx.download(f, n).view(np.uint32), ref.download(f, n).view(np.uint32))
        for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO)
    )
    log(f"cpu oracle: {k} steps in {dt:.2f}s on {cores} threads (max {max_threads()}); bitwise equal to GPU: {same}")
    if not same:
        raise RuntimeError("GPU and FP32 oracle differ on the benchmark mesh")
    return {
        "value": k * n / dt,
        "unit": "cell-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{k} Heun steps of the same {n}-cell mesh from the GPU state after the timed run "
        f"(FP32 C++/OpenMP oracle; bitwise equal to the GPU on those steps)",
    }


if __name__ == "__main__":
    sys.exit(main())
