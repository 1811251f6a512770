#!/usr/bin/env python3
"""Benchmark of the HIP resident Heun step on one MI355X.

Workload: the 16 003 008-cell rainfall flood (mixed triangle/quad lattice over
8 km of fBm terrain, 50 mm/h uniform rain from a dry start, per-cell Manning
n, FP32 with compensated depth, active set on). A "step" is one Heun step of
the whole mesh. After 20 warm-up steps the flood restarts from dry, runs 100
untimed lead steps through the wetting transient, then three consecutive
timed windows of 2000 steps; ``value`` comes from the median window. A
fourth window of the same flood times each stage kernel launch with HIP
events.

- ``value``: cell-steps/s over the timed steps, all cells counted every step
  (the reference's convention, flood_propagation.py:645-646), with the state
  in HBM before the timed region starts;
- ``roofline``: the corrector stage kernel (the larger of the two stage
  kernels), algorithmic bytes of the active blocks (28 B per edge, 4 B per
  CSR slot, 60 B per cell with the compensated low words) over its
  HIP-event-timed duration (fourth window), against 8 TB/s, with the HBM
  traffic measured by rocprofv3 counters and the attainable copy rates of
  the same box;
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
# HBM bytes per corrector launch on the 16 M-cell rainfall with every block
# active, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each
# scaled by what it reads for the library's plain copy of 1 GiB (FETCH x
# 2.0000, WRITE x 1.0000): mean over the 900 corrector launches of three
# 300-step windows, steps 101..1000 of this bench's flood.
CORRECTOR_TRAFFIC: dict | None = {
    "bytes_per_launch": (711391.2 * (1048576.0 / 524296.5) + 252588.2) * 1024.0,
    "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB plain copy; steps 101..1000, all blocks "
    "active",
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
    ap.add_argument("--lead", type=int, default=100, help="untimed steps from the dry start before the timed window")
    ap.add_argument("--no-legs", action="store_true", help="skip the compact batched and adjoint legs")
    ap.add_argument("--scale", type=int, default=1, help="shrink the lattice by this factor (testing)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-unordered", action="store_true", help="skip the unordered-accumulation comparison run")
    ap.add_argument(
        "--workload",
        choices=("rainfall", "batched", "adjoint", "inflow"),
        default="rainfall",
        help="rainfall: the plain 16 M-cell step (default); batched: scenarios on five Voronoi meshes, batched "
        "against sequential; adjoint: Manning/bed gradients of a gauge misfit on a 250 k-cell Voronoi flood; "
        "inflow: the 1 M-cell Voronoi inflow flood through the workflow, 60 one-minute hydrograph phases",
    )
    ap.add_argument("--cells", type=int, default=250_000, help="adjoint: Voronoi points")
    ap.add_argument("--fd-cells", type=int, default=20, help="adjoint: cells checked by central differences")
    ap.add_argument("--adj-thin", type=float, default=None,
                    help="adjoint: face depth (m) below which the reverse pass holds velocities (default 100 dry_tol)")
    ap.add_argument("--scenarios", type=int, default=64, help="batched: number of scenarios")
    ap.add_argument("--t-end", type=float, default=3600.0, help="batched: simulated seconds per scenario")
    ap.add_argument("--sizes", type=str, default="50000,200000,500000,1000000,2000000",
                    help="batched: Voronoi mesh sizes")
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
            "(it would shard by contiguous Hilbert block ranges, one halo exchange per stage); run with --gpus 1",
            file=sys.stderr,
        )
        return 2

    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "inflow":
        return bench_inflow(args)

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

    # Warm-up from the dry start (graphs recorded, clocks up), then the run
    # again from dry, untimed, through the wetting transient: the rain needs
    # about 15 steps to lift the first cells over dry_tol, and the thin film
    # of the next ~80 steps takes the costlier film branches. From step 100
    # on, the step time is flat (within 1 %) to step 2000, so a window that
    # starts `lead` = 100 steps in measures the run's steady step.
    start(ctx, wl, t_stop)
    if args.warmup:
        ctx.steps_timed(args.warmup)
    start(ctx, wl, t_stop)
    if args.lead:
        ctx.steps(args.lead, use_graph=True)
    lead_now = int(ctx.ctrl_read().now_ticks)
    log(f"lead: {args.lead} untimed steps from dry to t = {lead_now / 2**20:.2f} s")

    # The reference's protocol: the median of three timed runs after a
    # warm-up. The runs are three consecutive windows of the same flood
    # (steps lead+1 .. lead+3K), with no state upload between them, so the
    # GPU stays busy from the lead into the timed steps. A window's steps are
    # launched one by one with no event between them (as fast as replaying
    # the library's step graphs, and traceable by rocprofv3, whose kernel
    # trace aborts on graphs replayed back to back), synchronised on both
    # sides.
    runs = []
    for r in range(3):
        ctx.activity(reset=True)
        t0 = time.perf_counter()
        ctx.steps(args.steps, use_graph=False)
        wall = time.perf_counter() - t0
        act = ctx.activity()
        if act["steps"] != args.steps:
            raise RuntimeError(f"{args.steps - act['steps']} of the timed steps were no-ops")
        runs.append({"wall": wall, "act": act, "first_step": args.lead + r * args.steps + 1})
        log(f"run {r}: wall {wall:.3f}s ({wall / args.steps * 1e3:.4f} ms/step)")
    order = sorted(range(3), key=lambda k: runs[k]["wall"])
    med = runs[order[1]]
    wall, act = med["wall"], med["act"]
    # Each stage kernel's launch time: HIP events on the solver stream around
    # every predictor and corrector launch, over a fourth window of the same
    # flood (steps lead+3K+1 .. lead+4K). An event record between two launches
    # adds a few microseconds to the step, so these windows are not the ones
    # `value` is taken over.
    ctx.activity(reset=True)
    ms_pred, ms_corr, _ = ctx.steps_timed(args.steps)
    act_ev = ctx.activity()
    if act_ev["steps"] != args.steps:
        raise RuntimeError(f"{args.steps - act_ev['steps']} of the event-timed steps were no-ops")
    log(f"event window: predictor {ms_pred / args.steps:.4f} ms, corrector {ms_corr / args.steps:.4f} ms per launch")
    ctrl = ctx.ctrl_read()
    n, e, s = geom.N, geom.E, int(m["cptr"][-1])
    value = args.steps * n / wall
    # Algorithmic bytes (FP32/int32, each array once per stage): per stage
    # 28 B/edge + 4 B/slot + 4 B/cell CSR pointer + 8 B/cell statics + 12 B/cell
    # stage input + 12 B/cell output; the corrector also reads the start state
    # (12 B/cell); compensated low words 8 B/cell (predictor) and 12 B/cell
    # (corrector). Counted over active blocks only; uniform rain adds nothing.
    corr_bytes = 28 * act_ev["edges"] + 4 * act_ev["slots"] + (36 + 12 + 12) * act_ev["cells"]
    heun_bytes = step_bytes(act)
    corr_gbs = corr_bytes / (ms_corr * 1e-3) / 1e9
    step_gbs = step_bytes(act_ev) / ((ms_pred + ms_corr) * 1e-3) / 1e9
    active_frac = act["cells"] / (args.steps * n)
    log(f"median run: {wall / args.steps * 1e3:.4f} ms/step; active {active_frac:.3f}; corrector {corr_gbs:.0f} GB/s; "
        f"t_sim {ctrl.now_ticks / 2**20:.1f}s")

    # Attainable HBM rate: the library's plain copy kernel (also the HBM
    # counters' calibration kernel), tuned probes and hipMemcpy, 2 GiB each.
    nbw = 2 << 30
    copy_ms = ctx.stream_copy(nbw, 5)
    copy_gbs = 2 * nbw / (copy_ms * 1e-3) / 1e9
    bw = {"plain_copy_GBs": copy_gbs}
    best_copy = 0.0
    for unroll, wg in ((4, 8), (8, 4), (2, 8), (4, 4)):
        ms = ctx.stream_bw(nbw, 0, unroll, wg)
        rate = 2 * nbw / (ms * 1e-3) / 1e9
        bw[f"copy_u{unroll}_wg{wg}_GBs"] = rate
        best_copy = max(best_copy, rate)
    bw["read_only_GBs"] = nbw / (ctx.stream_bw(nbw, 1, 4, 8) * 1e-3) / 1e9
    bw["write_only_GBs"] = nbw / (ctx.stream_bw(nbw, 2, 4, 8) * 1e-3) / 1e9
    bw["hipMemcpy_d2d_GBs"] = 2 * nbw / (ctx.stream_bw(nbw, 3) * 1e-3) / 1e9
    best_copy = max(best_copy, bw["hipMemcpy_d2d_GBs"], copy_gbs)
    bw["best_copy_GBs"] = best_copy
    log(f"HBM probes (GB/s): {json.dumps({k: round(v) for k, v in bw.items()})}")
    if args.copy_calibration:
        ms = ctx.stream_copy(args.copy_calibration << 20, 3)
        log(f"copy calibration: 4 launches of k_stream_copy over {args.copy_calibration} MiB, {ms:.3f} ms each")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, wl, m, info, args.cpu_seconds)

    # The same window with unordered accumulation (INU_FLAG_DETERMINISTIC off,
    # the reference's atomic variant): its speed, and how far it drifts from
    # the fixed-order state over the window.
    unordered = None
    if not args.no_unordered:
        end_state = tuple(ctx.download(f, n) for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO))
        end_now = int(ctx.ctrl_read().now_ticks)
        start(ctx, wl, t_stop, state=end_state, now_ticks=end_now)
        ctx.steps_timed(args.steps)
        h_fixed = ctx.download(L.FIELD_H, n)
        ctx.set_params(9.81, wl.dry_tol, wl.cfl, flags & ~L.FLAG_DETERMINISTIC)
        start(ctx, wl, t_stop, state=end_state, now_ticks=end_now)
        ctx.steps_timed(min(args.steps, 20))
        start(ctx, wl, t_stop, state=end_state, now_ticks=end_now)
        t0 = time.perf_counter()
        u_pred, u_corr, _ = ctx.steps_timed(args.steps)
        u_wall = time.perf_counter() - t0
        h_u = ctx.download(L.FIELD_H, n)
        ctx.set_params(9.81, wl.dry_tol, wl.cfl, flags)
        unordered = {
            "ms_per_step": u_wall / args.steps * 1e3,
            "predictor_ms": u_pred / args.steps,
            "corrector_ms": u_corr / args.steps,
            "value": args.steps * n / u_wall,
            "max_abs_depth_diff_m": float(np.abs(h_u.astype(np.float64) - h_fixed).max()),
        }
        log(f"unordered accumulation: {unordered}")
    ctx.close()
    del ctx

    legs = {}
    if not args.no_legs:
        legs["batched"] = batched_leg(heun_bytes / wall)
        legs["adjoint"] = adjoint_leg(args)

    # The measured traffic belongs to the full-size lattice only.
    traffic = CORRECTOR_TRAFFIC if args.scale == 1 else None
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
            "window": f"three windows of {args.steps} steps, steps {args.lead + 1}..{args.lead + 3 * args.steps} "
            f"of the run from dry, launched step by step; the median one starts at step "
            f"{med['first_step']}; then {args.steps} steps with HIP events around each stage kernel (to t = "
            f"{ctrl.now_ticks / 2**20:.1f} s, from t = {lead_now / 2**20:.1f} s after the lead)",
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
            "traffic": None if traffic is None else traffic["bytes_per_launch"],
            "traffic_over_algorithmic": None if traffic is None
            else traffic["bytes_per_launch"] / (corr_bytes / args.steps),
            "traffic_source": None if traffic is None else traffic["source"],
            "ms_per_launch": ms_corr / args.steps,
            "ms_per_launch_is": f"HIP events on the solver stream around each corrector launch, steps "
            f"{args.lead + 3 * args.steps + 1}..{args.lead + 4 * args.steps}",
            "algorithmic_bytes_per_launch": corr_bytes / args.steps,
            "frac_of_best_copy": corr_gbs / bw["best_copy_GBs"],
            "hbm_probes_GBs": bw,
        },
        "heun_step_roofline": {
            "achieved": step_gbs,
            "frac": step_gbs / HBM_PEAK_GBS,
            "unit": "GB/s",
            "bytes_per_step": heun_bytes / args.steps,
            "stage_kernels_ms_per_step": (ms_pred + ms_corr) / args.steps,
            # The same bytes over the whole step's wall time: the stage
            # kernels, the per-step clock, dilation and source kernels and
            # the gaps between launches.
            "whole_step_achieved": heun_bytes / wall / 1e9,
            "whole_step_frac": heun_bytes / wall / 1e9 / HBM_PEAK_GBS,
            "active_cell_steps_per_s_over_stage_kernels": act["cells"] / ((ms_pred + ms_corr) * 1e-3),
        },
        "median_of_3": {
            "ms_per_step": [r["wall"] / args.steps * 1e3 for r in runs],
            "value_is": "the median run",
        },
        "cpu_baseline": cpu,
        "unordered_accumulation": unordered,
        **legs,
    }
    print(json.dumps(out))
    return 0


def _phase_loop(ctx, geoms, scen, n_phases: int, block: int, first: int = 0, verbose: bool = False) -> None:  # noqa: ANN001
    """Run every scenario of ctx through its one-minute phases, each scenario on its own schedule.

    A scenario that reaches the end of a phase moves on at once: its inflow
    rate for the next minute, its stop one minute on and a retarget. The
    others keep stepping; a retarget leaves a scenario whose stop did not
    move bit for bit where it was (the same next dt from its last bound,
    the same lists). Alone, this is the plain phase loop.
    """
    sel = scen[first:first + ctx.n_scenarios]
    base = np.cumsum([0] + [geoms[s.mesh].N for s in sel])
    pairs = [s.source_pairs(geoms[s.mesh], 30.0) for s in sel]
    cells = np.concatenate([c + np.uint32(base[k]) for k, (c, _) in enumerate(pairs)]).astype(np.uint32)
    ctx.set_sources(cells, np.concatenate([r for _, r in pairs]).astype(np.float32))
    cs = ctx.ctrls_read()
    for c in cs:
        c.stop_ticks = int(c.stop_ticks) + 60 * 2**20
    ctx.ctrl_write(cs)
    ctx.prologue()
    phase = [0] * len(sel)
    reported = 0
    while True:
        ctx.steps(block, use_graph=True)
        cs = ctx.ctrls_read()
        at_stop = [int(c.now_ticks) >= int(c.stop_ticks) for c in cs]
        move = [k for k in range(len(sel)) if at_stop[k] and phase[k] + 1 < n_phases]
        if not move:
            if all(at_stop):
                break
            continue
        # Only the inflow cells' rates change between phases.
        rates = []
        for k in move:
            phase[k] += 1
            c_k, r_k = sel[k].source_pairs(geoms[sel[k].mesh], (phase[k] + 0.5) * 60.0)
            rates.append(r_k[int(np.searchsorted(c_k, sel[k].inflow_cell))])
            cs[k].stop_ticks = int(cs[k].stop_ticks) + 60 * 2**20
        ctx.set_source_rates(np.array([base[k] + sel[k].inflow_cell for k in move], np.uint32),
                             np.array(rates, np.float32))
        ctx.ctrl_write(cs)
        ctx.retarget()
        if verbose and min(phase) > reported:
            reported = min(phase)
            log(f"every scenario in phase {reported + 1}/{n_phases} or later")
    if verbose:
        log(f"all {len(sel)} scenarios through {n_phases} phases")


def _start_ctrls(n: int):  # noqa: ANN202
    from inundation_amd.hip.lib import InuCtrl
    from inundation_amd.tuning import dt_to_ticks

    out = []
    for _ in range(n):
        c = InuCtrl()
        c.now_ticks = 0
        c.stop_ticks = 0
        c.steps = 0
        c.dt_ticks = 0
        c.dt_max_ticks = dt_to_ticks(1.0)
        c.dry_dt_ticks = dt_to_ticks(0.1)
        c.cfl_safety = 1.0
        out.append(c)
    return out


def step_bytes(act: dict) -> int:
    """Algorithmic bytes of both stage kernels over the activity counters' cells, edges and slots."""
    return 56 * act["edges"] + 8 * act["slots"] + 104 * act["cells"]


def lattice_fill_rate(lead: int = 100, steps: int = 300) -> float:
    """Algorithmic bytes per second of whole Heun steps on the 16 M-cell rainfall, every block active.

    The rate the stage kernels reach on a mesh that fills the GPU: steps
    lead+1 .. lead+steps of the rainfall flood from dry, replayed as one
    hipGraph, by the wall clock.
    """
    from inundation_amd.hip import lib as L
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workloads import rainfall

    wl = rainfall()
    m = device_mesh_arrays(wl.geom)
    ctx = L.Context(0)
    ctx.set_mesh(m["cell_l"], m["cell_r"], m["nx"], m["ny"], m["len"], m["dz2"], m["area"], m["cptr"], m["cidx"])
    ctx.set_manning(wl.manning)
    ctx.set_params(9.81, wl.dry_tol, wl.cfl, L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET)
    ctx.set_sources(*wl.source_pairs())
    start(ctx, wl, 1800.0)
    ctx.steps(lead, use_graph=True)
    ctx.activity(reset=True)
    t0 = time.perf_counter()
    ctx.steps(steps, use_graph=True)
    wall = time.perf_counter() - t0
    rate = step_bytes(ctx.activity()) / wall
    ctx.close()
    log(f"fill rate: the 16 M-cell rainfall moves {rate / 1e9:.0f} GB/s of algorithmic bytes")
    return rate


def run_batched(sizes: tuple[int, ...], n_scenarios: int, t_end: float, processes: int = 5,
                fill_Bps: float | None = None) -> dict:
    """Scenarios on five Voronoi meshes: batched in one context against one by one (the plain solver).

    Both runs take each scenario from a dry start through one-minute phases
    (the inflow hydrograph changes per phase, the rain is constant),
    replaying 50-step hipGraphs; a scenario moves to its next phase as soon
    as it reaches the end of one (``_phase_loop``). The wall clock covers the
    phase loops only; cell-steps count each scenario's cells times its real
    steps. Each batched scenario is bit-identical to it alone, so both runs
    take the same steps.

    The bound on the speed-up: the algorithmic bytes of all the visited
    blocks' steps (the device's activity counters), moved at ``fill_Bps``,
    the rate the same kernels reach on a mesh that fills the GPU (measured
    on the 16 M-cell rainfall when not given), against the sequential time.
    """
    from inundation_amd.hip import lib as L
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workloads import batched

    t0 = time.perf_counter()
    geoms, scen = batched(n_scenarios=n_scenarios, sizes=sizes, progress=log, processes=processes)
    log(f"{len(geoms)} meshes {[g.N for g in geoms]} built in {time.perf_counter() - t0:.1f}s; {len(scen)} scenarios")
    dev = [device_mesh_arrays(g) for g in geoms]
    n_phases = max(1, int(round(t_end / 60.0)))
    flags = L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET
    block = 50

    def setup(ctx, sel) -> None:  # noqa: ANN001
        n_all = sum(geoms[s.mesh].N for s in sel)
        ctx.set_params(9.81, 1e-4, 0.8, flags)
        ctx.set_manning(np.full(n_all, 0.035, np.float32))
        z = np.zeros(n_all, np.float32)
        ctx.set_state(z, z, z, None)
        ctx.ctrl_write(_start_ctrls(len(sel)))

    def autotune_on_flood(ctx, sel) -> dict:  # noqa: ANN001
        """Tile level from a state 100 steps into the sources of minute 0.5 (the state is reset afterwards)."""
        base_ = np.cumsum([0] + [geoms[s.mesh].N for s in sel])
        prs = [s.source_pairs(geoms[s.mesh], 30.0) for s in sel]
        ctx.set_sources(np.concatenate([c + np.uint32(base_[k]) for k, (c, _) in enumerate(prs)]).astype(np.uint32),
                        np.concatenate([r for _, r in prs]).astype(np.float32))
        cs_ = ctx.ctrls_read()
        for c_ in cs_:
            c_.stop_ticks = 10**6 * 2**20
        ctx.ctrl_write(cs_)
        ctx.prologue()
        ctx.steps(100, use_graph=False)
        return ctx.autotune(5)

    # Sequential: the plain solver, one scenario at a time, one context per
    # mesh, each mesh at the tile level the autotuner picks for it.
    seq_wall = 0.0
    seq_bytes = 0
    seq_steps: dict[int, int] = {}
    alone: dict[int, dict] = {}
    for m, g in enumerate(geoms):
        ctx = L.Context(0)
        d = dev[m]
        ctx.set_mesh(d["cell_l"], d["cell_r"], d["nx"], d["ny"], d["len"], d["dz2"], d["area"], d["cptr"], d["cidx"])
        first = next(k for k, s in enumerate(scen) if s.mesh == m)
        setup(ctx, [scen[first]])
        tile_m = autotune_on_flood(ctx, [scen[first]])["tile"]
        mesh_wall, mesh_cs, mesh_bytes = 0.0, 0, 0
        for k, s in enumerate(scen):
            if s.mesh != m:
                continue
            setup(ctx, [s])
            ctx.activity(reset=True)
            t0 = time.perf_counter()
            _phase_loop(ctx, geoms, scen, n_phases, block, first=k)
            w = time.perf_counter() - t0
            seq_wall += w
            mesh_wall += w
            seq_steps[k] = int(ctx.ctrl_read().steps)
            mesh_cs += seq_steps[k] * g.N
            mesh_bytes += step_bytes(ctx.activity())
            log(f"sequential: scenario {k} on mesh {m}: {seq_steps[k]} steps, {seq_wall:.1f}s so far")
        ctx.close()
        seq_bytes += mesh_bytes
        alone[g.N] = {"tile": tile_m, "wall_s": mesh_wall, "cell_steps": mesh_cs, "value": mesh_cs / mesh_wall,
                      "algorithmic_GBs": mesh_bytes / mesh_wall / 1e9}
        log(f"sequential: mesh {m} ({g.N} cells, tile {tile_m}) done: {mesh_cs / mesh_wall / 1e9:.2f} G cell-steps/s "
            f"alone, {seq_wall:.1f}s so far")
    seq_cell_steps = sum(seq_steps[k] * geoms[s.mesh].N for k, s in enumerate(scen))

    # Batched: every scenario in one context, tile level autotuned on a
    # state 100 steps into the first phase, then a fresh start.
    ctx = L.Context(0)
    ctx.set_batch(dev, [s.mesh for s in scen])
    setup(ctx, scen)
    base = np.cumsum([0] + [geoms[s.mesh].N for s in scen])
    tune = autotune_on_flood(ctx, scen)
    log(f"autotune: tile {tune['tile']}, predictor us/launch by tile {tune['us_per_launch']}")
    setup(ctx, scen)
    ctx.activity(reset=True)
    t0 = time.perf_counter()
    _phase_loop(ctx, geoms, scen, n_phases, block, verbose=True)
    bat_wall = time.perf_counter() - t0
    bat_steps = [int(c.steps) for c in ctx.ctrls_read()]
    bat_bytes = step_bytes(ctx.activity())
    ctx.close()
    # Batching can at best move the visited blocks' bytes at the rate the
    # kernels reach on a mesh that fills the GPU.
    if fill_Bps is None:
        fill_Bps = lattice_fill_rate()
    bound = seq_wall / (seq_bytes / fill_Bps)
    bat_cell_steps = sum(st * geoms[s.mesh].N for st, s in zip(bat_steps, scen))
    same = bat_steps == [seq_steps[k] for k in range(len(scen))]
    seq_rate = seq_cell_steps / seq_wall
    bat_rate = bat_cell_steps / bat_wall
    log(f"sequential {seq_wall:.2f}s {seq_rate / 1e9:.2f} G cell-steps/s; batched {bat_wall:.2f}s "
        f"{bat_rate / 1e9:.2f} G cell-steps/s; speedup {bat_rate / seq_rate:.2f}x (bound {bound:.2f}x); "
        f"same steps: {same}; same bytes: {bat_bytes == seq_bytes}")
    if not same:
        raise RuntimeError("batched and sequential runs took different steps")
    out = {
        "metric": METRIC,
        "value": bat_rate,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": int(sum(bat_steps)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {len(scen)} scenarios on Voronoi meshes of {list(sizes)} cells, "
            f"{t_end:g} s simulated each, inflow U[5,50] m^3/s + rain U[0,80] mm/h",
            "scenarios": len(scen),
            "cells_total": int(base[-1]),
            "phases": n_phases,
            "graph_steps": block,
            "tile": tune["tile"],
            "tile_predictor_us": tune["us_per_launch"],
        },
        "batched": {"wall_s": bat_wall, "cell_steps": bat_cell_steps, "value": bat_rate,
                    "algorithmic_GBs": bat_bytes / bat_wall / 1e9},
        "sequential": {"wall_s": seq_wall, "cell_steps": seq_cell_steps, "value": seq_rate,
                       "algorithmic_GBs": seq_bytes / seq_wall / 1e9,
                       "tile_per_mesh": {str(k): v["tile"] for k, v in alone.items()}},
        "alone_by_mesh_cells": {str(k): v for k, v in alone.items()},
        "speedup": bat_rate / seq_rate,
        "speedup_bound": bound,
        "fill_GBs": fill_Bps / 1e9,
        "speedup_bound_is": "the sequential time over the time to move every visited block's algorithmic "
        "bytes at fill_GBs, the rate of the same kernels on the 16 M-cell rainfall (every block active)",
    }
    return out


def run_adjoint(cells: int, steps: int, fd_cells: int, cpu_seconds: float | None, thin: float | None = None) -> dict:
    """Gauge-misfit gradients w.r.t. Manning n and bed z: the adjoint against the plain forward run.

    A truth run with the smooth 'true' n field picks 32 gauges among the cells
    it wets and records their depths every 10 steps; the calibration run
    starts from n = 0.04. Timed: the plain forward (hipGraph steps, no
    recording), and the gradient (recorded forward + backward: segment replays
    from checkpoints every 100 steps and the reverse kernels). Central
    differences (eps = 1e-3) of replays of the recording (same dt sequence and
    source events) check the gradient on ``--fd-cells`` cells drawn at random
    among the cells the run wets, reported per cell with the spread between
    eps and eps/3 that marks a difference crossing a branch.
    """
    from inundation_amd.hip import lib as L
    from inundation_amd.hip.adjoint import GaugeRun
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workloads import calibration

    t0 = time.perf_counter()
    cal = calibration(n_points=cells, steps=steps)
    geom = cal.geom
    log(f"calibration mesh: {geom.N} cells built in {time.perf_counter() - t0:.1f}s")
    d = device_mesh_arrays(geom)
    ctx = L.Context(0)
    ctx.set_mesh(d["cell_l"], d["cell_r"], d["nx"], d["ny"], d["len"], d["dz2"], d["area"], d["cptr"], d["cidx"])
    zero = np.zeros(geom.N, np.float32)
    run = GaugeRun(ctx=ctx, state0=(zero, zero, zero), gauges=np.array([cal.inflow_cell], np.uint32),
                   every=cal.every, n_steps=cal.steps, seg=cal.seg, dry_tol=1e-4, cfl=0.8, dt_max=1.0, dt_init=0.1,
                   source_cells=cal.source_cells(), rates_at=cal.rates_at, graph_steps=50)
    if thin is not None:
        ctx.adj_set_thin(thin)
    z = geom.zb.astype(np.float32)
    # Truth run: gauges among the cells it wets, then their observed depths.
    run.forward_plain(cal.n_true, z)
    h_truth = ctx.download(L.FIELD_H, geom.N)
    run.gauges = cal.pick_gauges(h_truth)
    obs = run.forward(cal.n_true, z).astype(np.float64)
    t_truth = ctx.ctrl_read().now_ticks / float(1 << 20)
    log(f"truth run: {cal.steps} steps, {t_truth:.1f} s simulated, {int((h_truth > 1e-4).sum())} wet cells, "
        f"gauge depths {obs[-1].min():.3f}..{obs[-1].max():.3f} m")

    n0 = np.full(geom.N, cal.n_start, np.float32)
    reps = 3
    t_plain = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run.forward_plain(n0, z)
        t_plain.append(time.perf_counter() - t0)
    h_end = ctx.download(L.FIELD_H, geom.N)
    active_frac = float(ctx.download(L.FIELD_BLOCK_ACTIVE, -(-geom.N // 256), np.uint32).astype(bool).mean())
    t_fwd, t_bwd = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        series = run.forward(n0, z)
        t1 = time.perf_counter()
        seeds = (2.0 * (series.astype(np.float64) - obs)).astype(np.float32)
        gn, gz = run.backward(seeds)
        t2 = time.perf_counter()
        t_fwd.append(t1 - t0)
        t_bwd.append(t2 - t1)
    loss = float(((series.astype(np.float64) - obs) ** 2).sum())
    plain, fwd, bwd = min(t_plain), min(t_fwd), min(t_bwd)
    ratio = (fwd + bwd) / plain
    log(f"plain forward {plain * 1e3:.1f} ms, recorded forward {fwd * 1e3:.1f} ms, backward {bwd * 1e3:.1f} ms: "
        f"gradient = {ratio:.2f}x forward; loss {loss:.4e}; wet cells at the end {int((h_end > 1e-4).sum())}")

    # The same gradient under binomial checkpointing: 5 checkpoint slots
    # instead of one per segment; the extra replays cost time, not bits.
    slots = 5
    run.slots = slots
    t_bfwd, t_bbwd = [], []
    try:
        for _ in range(2):
            t0 = time.perf_counter()
            run.forward(n0, z)
            t1 = time.perf_counter()
            gn_b, gz_b = run.backward(seeds)
            t2 = time.perf_counter()
            t_bfwd.append(t1 - t0)
            t_bbwd.append(t2 - t1)
        replayed = ctx.adj_replayed()
    finally:
        run.slots = 0
    n_seg = -(-cal.steps // cal.seg)
    binomial = {
        "slots": slots,
        "segments": n_seg,
        "checkpoint_bytes": slots * 16 * geom.N,
        "uniform_checkpoint_bytes": (n_seg + 1) * 16 * geom.N,
        "recorded_forward_s": min(t_bfwd),
        "backward_s": min(t_bbwd),
        "cost_over_forward": (min(t_bfwd) + min(t_bbwd)) / plain,
        "steps_advanced_without_tape": replayed["advanced"],
        "steps_replayed_with_tape": replayed["taped"],
        "gradient_bitwise_uniform": bool(np.array_equal(gn_b.view(np.uint32), gn.view(np.uint32))
                                         and np.array_equal(gz_b.view(np.uint32), gz.view(np.uint32))),
    }
    log(f"binomial checkpointing: {binomial}")

    # Central differences of replays (the recorded dt sequence and source events).
    def loss_of(nn, zz) -> float:  # noqa: ANN001
        s_ = run.replay(nn, zz).astype(np.float64)
        return float(((s_ - obs) ** 2).sum())

    replay_same = np.array_equal(run.replay(n0, z).view(np.uint32), series.view(np.uint32))
    rng = np.random.default_rng([cal.seed, 2])
    fd = {}
    eps, eps2 = 1e-3, 3e-4
    # The configuration's protocol: random cells, here among the cells the
    # run wets (elsewhere both gradient and difference are zero).
    cand = np.flatnonzero(h_end > run.dry_tol)
    for name, vec, grad in (("n", n0, gn), ("z", z, gz)):
        cells = np.sort(rng.choice(cand, min(fd_cells, cand.size), replace=False))
        rows = []
        for i in cells.tolist():
            f = []
            for e in (eps, eps2):
                up, dn = vec.copy(), vec.copy()
                up[i] += e
                dn[i] -= e
                if name == "n":
                    f.append((loss_of(up, z) - loss_of(dn, z)) / (2 * e))
                else:
                    f.append((loss_of(n0, up) - loss_of(n0, dn)) / (2 * e))
            rows.append((i, float(grad[i]), f[0], f[1]))
        g_ = np.array([r[1] for r in rows])
        f_ = np.array([r[2] for r in rows])
        f2 = np.array([r[3] for r in rows])
        rel = np.abs(g_ - f_) / np.maximum(np.abs(f_), 1e-30)
        # A difference that changes by more than 1 % between eps and eps/3
        # crosses a branch (a wet/dry or wave-speed switch moves inside the
        # interval): the adjoint follows the forward's branches, so there the
        # difference is not the derivative it computes.
        spread = np.abs(f2 - f_) / np.maximum(np.abs(f_), 1e-30)
        smooth = spread < 1e-2
        fd[name] = {
            "cells": int(cells.size),
            "drawn_from": f"{cand.size} cells wet at the end of the run, at random",
            "rel_err_norm": float(np.linalg.norm(g_ - f_) / max(np.linalg.norm(f_), 1e-30)),
            "rel_err_median": float(np.median(rel)),
            "rel_err_max": float(rel.max()),
            "fd_eps_spread_norm": float(np.linalg.norm(f2 - f_) / max(np.linalg.norm(f_), 1e-30)),
            "branch_consistent_cells": int(smooth.sum()),
            "rel_err_norm_branch_consistent": float(np.linalg.norm(g_[smooth] - f_[smooth])
                                                    / max(np.linalg.norm(f_[smooth]), 1e-30)) if smooth.any() else None,
            "per_cell": [{"cell": int(r[0]), "adjoint": r[1], "fd": r[2], "fd_eps3": r[3], "rel_err": float(e),
                          "fd_spread": float(sp)} for r, e, sp in zip(rows, rel, spread)],
        }
        log(f"FD {name}: { {k: v for k, v in fd[name].items() if k != 'per_cell'} }")
        for r in rows[:5]:
            log(f"   cell {r[0]}: adjoint {r[1]:.6e} fd(1e-3) {r[2]:.6e} fd(3e-4) {r[3]:.6e}")
    stats = ctx.adj_stats()
    cpu = None if cpu_seconds is None else cpu_baseline_adjoint(ctx, cal, run, n0, z, cpu_seconds)
    ctx.close()
    wet = adjoint_cost_wet()
    cell_steps = geom.N * cal.steps
    out = {
        "metric": METRIC,
        "value": cell_steps / (fwd + bwd),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": cal.steps,
        "warmup": 1,
        "ms_per_step": (fwd + bwd) / cal.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"adjoint calibration: {geom.N} Voronoi cells over 2 km fBm with a carved channel, "
            f"20 m^3/s-peak inflow hydrograph, true n smooth in [0.02,0.08], start n=0.04, {cal.n_gauges} gauges "
            f"every {cal.every} steps, {cal.steps} steps, checkpoints every {cal.seg}",
            "cells": geom.N,
            "simulated_s": t_truth,
            "value_is": "cell-steps of the calibration run differentiated per second (recorded forward + backward)",
        },
        "adjoint": {
            "plain_forward_s": plain,
            "recorded_forward_s": fwd,
            "backward_s": bwd,
            "cost_over_forward": ratio,
            "loss": loss,
            "replay_bitwise": bool(replay_same),
            "binomial": binomial,
            "active_block_fraction_at_end": active_frac,
            "fd_eps": [eps, eps2],
            "fd": fd,
            "thin_layer_depth_m": 100.0 * run.dry_tol if thin is None else thin,
            "reverse_stats": stats,
            "wet_mesh_cost": wet,
        },
        "cpu_baseline": cpu,
    }
    return out


def bench_inflow(args) -> int:  # noqa: ANN001
    """The 1 M-cell Voronoi inflow flood, 1 h simulated, through ``SWEWorkflow`` on the hip_resident backend.

    60 one-minute phases carry the triangular hydrograph (peak 20 m^3/s at
    30 min) into one source cell; outputs every 60 s. The value is the
    reference's convention: steps x cells over the run loop's wall clock
    (``WorkflowResult.wall_seconds``, snapshot downloads included). The CPU
    baseline is the FP32 oracle continuing from the GPU's final state for a
    bounded number of steps, checked bit for bit against the GPU on them.
    """
    from inundation_amd.hip import lib as L
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workflow import SWEWorkflow, WorkflowConfig
    from inundation_amd.workloads import inflow

    t0 = time.perf_counter()
    wl, phases = inflow()
    geom = wl.geom
    log(f"mesh {geom.N} cells, {geom.E} edges built in {time.perf_counter() - t0:.1f}s")
    cfg = WorkflowConfig(mesh_source="synthetic", manning_n=0.035, output_interval_s=60.0, dt_max=wl.dt_max,
                         dt_init=wl.dt_init, cfl_interval=10, dry_tol=wl.dry_tol, cfl=wl.cfl, progress=False)
    runs = []
    for r in range(2):  # a warm-up run, then the timed one
        wf = SWEWorkflow(cfg)
        wf.prepare(geometry=geom)
        res = wf.run(phases)
        runs.append(res)
        log(f"run {r}: {res.steps_total} steps in {res.wall_seconds:.2f}s, {len(res.snapshots)} snapshots, "
            f"volume {res.volume_final_m3:.1f} of {res.volume_injected_m3:.1f} m^3 injected")
        if r == 0:
            wf.solver.close()
    res = runs[-1]
    if [np.asarray(a).view(np.uint32).tobytes() for a in runs[0].snapshots] != [
            np.asarray(a).view(np.uint32).tobytes() for a in res.snapshots]:
        raise RuntimeError("the two runs differ")
    value = res.steps_total * geom.N / res.wall_seconds
    ctx = wf.solver.context
    wet = int((res.h_final >= wl.dry_tol).sum())
    cpu = None
    if not args.no_cpu:
        from oracle.cpu_ref import CpuRef, set_threads

        m = device_mesh_arrays(geom)
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:  # pragma: no cover
            cores = os.cpu_count() or 1
        cores = max(1, min(cores, 16))
        set_threads(cores)
        ref = CpuRef(m)
        state = tuple(ctx.download(f, geom.N) for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO))
        ctrl = ctx.ctrl_read()
        ref.set_params(9.81, wl.dry_tol, wl.cfl, L.FLAG_DETERMINISTIC | L.FLAG_COMPENSATED | L.FLAG_ACTIVE_SET)
        ref.set_manning(wl.manning)
        ref.set_sources(np.empty(0, np.uint32), np.empty(0, np.float32))
        ctx.set_sources(np.empty(0, np.uint32), np.empty(0, np.float32))
        for be in (ref, ctx):
            be.set_state(*state)
            c = L.InuCtrl.from_buffer_copy(ctrl)
            c.steps = 0
            c.stop_ticks = int(c.now_ticks) + 3600 * 2**20
            be.ctrl_write(c)
            be.prologue()
        t1 = time.perf_counter()
        ref.steps(1)
        one = time.perf_counter() - t1
        k = max(1, min(2000, int(args.cpu_seconds / max(one, 1e-6)) - 1))
        t1 = time.perf_counter()
        ref.steps(k)
        dt = time.perf_counter() - t1
        ctx.steps(k + 1, use_graph=False)
        same = all(np.array_equal(ctx.download(f, geom.N).view(np.uint32), ref.download(f, geom.N).view(np.uint32))
                   for f in (L.FIELD_H, L.FIELD_HU, L.FIELD_HV, L.FIELD_H_LO))
        log(f"cpu oracle: {k} steps in {dt:.2f}s on {cores} threads; bitwise equal to GPU: {same}")
        if not same:
            raise RuntimeError("GPU and FP32 oracle differ on the inflow mesh")
        cpu = {"value": k * geom.N / dt, "unit": "cell-steps/s", "cores": cores, "kind": "port",
               "sample": f"{k} Heun steps of the same {geom.N}-cell mesh from the GPU state after the hour (FP32 "
               f"C++/OpenMP oracle, active set on; bitwise equal to the GPU on those steps)"}
    wf.solver.close()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": int(res.steps_total),
        "warmup": 1,
        "ms_per_step": res.wall_seconds / res.steps_total * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl.name + ", 60 one-minute phases, outputs every 60 s, through SWEWorkflow",
            "cells": geom.N,
            "edges": geom.E,
            "simulated_s": 3600.0,
            "wall_s": res.wall_seconds,
            "snapshots": len(res.snapshots),
            "wet_cells_at_end": wet,
            "volume_final_m3": res.volume_final_m3,
            "volume_injected_m3": res.volume_injected_m3,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def bench_batched(args) -> int:  # noqa: ANN001
    sizes = tuple(int(x) for x in args.sizes.split(","))
    print(json.dumps(run_batched(sizes, args.scenarios, args.t_end)))
    return 0


def bench_adjoint(args) -> int:  # noqa: ANN001
    print(json.dumps(run_adjoint(args.cells, args.steps, args.fd_cells, None if args.no_cpu else args.cpu_seconds,
                                 args.adj_thin)))
    return 0


def batched_leg(fill_Bps: float) -> dict:
    """The batched leg of the metric, compact: 10 scenarios on 50 k- and 200 k-cell Voronoi meshes, 5 minutes each.

    ``fill_Bps``: the main run's algorithmic bytes per second, the bound's rate.
    """
    r = run_batched((50_000, 200_000), 10, 300.0, processes=1, fill_Bps=fill_Bps)
    keep = ("batched", "sequential", "speedup", "speedup_bound", "fill_GBs", "alone_by_mesh_cells")
    return {"workload": r["config"]["workload"], "tile": r["config"]["tile"], **{k: r[k] for k in keep}}


def adjoint_leg(args) -> dict:  # noqa: ANN001
    """The adjoint leg of the metric: the 250 k-cell calibration gradient, 2000 steps, 5 cells checked by
    central differences (the full check is ``--workload adjoint``)."""
    del args
    r = run_adjoint(250_000, 2000, 5, None)
    a = r["adjoint"]
    return {"workload": r["config"]["workload"], "value": r["value"], "unit": r["unit"],
            "cost_over_forward": a["cost_over_forward"], "plain_forward_s": a["plain_forward_s"],
            "recorded_forward_s": a["recorded_forward_s"], "backward_s": a["backward_s"], "fd": a["fd"],
            "replay_bitwise": a["replay_bitwise"], "binomial": a["binomial"]}


def adjoint_cost_wet(scale: int = 2, steps: int = 300, seg: int = 50) -> dict:
    """Cost of the gradient on a mesh the flood covers: the rainfall lattice at 1/scale (about 4 M cells).

    300 steps from dry under 50 mm/h on every cell (wet from step 15 on), 32
    random gauges every 10 steps, checkpoints every ``seg`` steps; the
    plain forward against the recorded forward plus the backward, each the
    best of two. Device memory: the tape holds seg x 56 bytes per cell.
    """
    from inundation_amd.hip import lib as L
    from inundation_amd.hip.adjoint import GaugeRun
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workloads import rainfall

    wl = rainfall(scale=scale)
    geom = wl.geom
    d = device_mesh_arrays(geom)
    ctx = L.Context(0)
    ctx.set_mesh(d["cell_l"], d["cell_r"], d["nx"], d["ny"], d["len"], d["dz2"], d["area"], d["cptr"], d["cidx"])
    cells, rates = wl.source_pairs()
    gauges = np.sort(np.random.default_rng([wl.seed, 3]).choice(geom.N, 32, replace=False)).astype(np.uint32)
    z0 = np.zeros(geom.N, np.float32)
    run = GaugeRun(ctx=ctx, state0=(z0, z0, z0), gauges=gauges, every=10, n_steps=steps, seg=seg, dry_tol=wl.dry_tol,
                   cfl=wl.cfl, dt_max=wl.dt_max, dt_init=wl.dt_init, source_cells=cells, phases=[(0, rates)],
                   graph_steps=50)
    z = geom.zb.astype(np.float32)
    t_plain, t_fwd, t_bwd = [], [], []
    for _ in range(2):
        t0 = time.perf_counter()
        run.forward_plain(wl.manning, z)
        t_plain.append(time.perf_counter() - t0)
    active = float(ctx.download(L.FIELD_BLOCK_ACTIVE, -(-geom.N // 256), np.uint32).astype(bool).mean())
    for _ in range(2):
        t0 = time.perf_counter()
        series = run.forward(wl.manning, z)
        t1 = time.perf_counter()
        gn, gz = run.backward((2.0 * series.astype(np.float64)).astype(np.float32))
        t_fwd.append(t1 - t0)
        t_bwd.append(time.perf_counter() - t1)
    ctx.close()
    plain, fwd, bwd = min(t_plain), min(t_fwd), min(t_bwd)
    out = {"cells": geom.N, "steps": steps, "seg": seg, "active_block_fraction_at_end": active,
           "plain_forward_s": plain, "recorded_forward_s": fwd, "backward_s": bwd,
           "cost_over_forward": (fwd + bwd) / plain, "tape_bytes": seg * 56 * geom.N,
           "gradient_nonzero_cells": {"n": int(np.count_nonzero(gn)), "z": int(np.count_nonzero(gz))}}
    log(f"adjoint on the wet {geom.N}-cell rainfall mesh: {out}")
    return out


def cpu_baseline_adjoint(ctx, cal, run, n0, z, budget_s: float) -> dict:  # noqa: ANN001
    """The FP64 oracle's reverse-mode gradient (torch autodiff on the host cores) on a bounded slice.

    The slice is the calibration run's first K steps from the dry start (the
    recorded dt and source rate), with the loss taken at the 32 cells nearest
    the inlet every 5 steps (sum of squared depths) so that it sees water this
    early. The same slice runs through the GPU adjoint, and the two gradients
    are compared. About 1.7 GB of host memory per step at 250 k cells.
    """
    import torch

    from inundation_amd.hip.adjoint import GaugeRun
    from inundation_amd.hip.solver import device_mesh_arrays
    from oracle import adjoint64 as A

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    torch.set_num_threads(cores)
    geom = cal.geom
    k_steps = 20 if budget_s >= 10.0 else 10
    every = 5
    d2 = (geom.centroid[:, 0] - geom.centroid[cal.inflow_cell, 0]) ** 2 + (
        geom.centroid[:, 1] - geom.centroid[cal.inflow_cell, 1]
    ) ** 2
    gauges = np.sort(np.argsort(d2)[:32]).astype(np.uint32)
    zero = np.zeros(geom.N, np.float32)
    sl = GaugeRun(ctx=ctx, state0=(zero, zero, zero), gauges=gauges, every=every, n_steps=k_steps, seg=10,
                  dry_tol=run.dry_tol, cfl=run.cfl, dt_max=run.dt_max, dt_init=run.dt_init,
                  source_cells=cal.source_cells(), rates_at=cal.rates_at, graph_steps=k_steps)
    series = sl.forward(n0, z)
    gn, gz = sl.backward((2.0 * series.astype(np.float64)).astype(np.float32))
    dts = sl.dts().astype(np.float64)
    t_end = float(dts.sum())
    if t_end >= 60.0:
        raise RuntimeError("the CPU slice crosses a hydrograph phase")
    src = np.zeros(geom.N)
    src[cal.inflow_cell] = float(cal.rates_at(0.0)[0])
    m = device_mesh_arrays(geom)
    t0 = time.perf_counter()
    _, gz64, gn64 = A.misfit_and_grads(m, z.astype(np.float64), n0.astype(np.float64),
                                       tuple(np.zeros(geom.N) for _ in range(3)), dts, src, gauges,
                                       np.zeros((k_steps // every, gauges.size)), every, thin=100.0 * run.dry_tol)
    dt = time.perf_counter() - t0
    err_n = float(np.linalg.norm(gn - gn64) / np.linalg.norm(gn64))
    err_z = float(np.linalg.norm(gz - gz64) / np.linalg.norm(gz64))
    log(f"cpu FP64 adjoint: {k_steps} steps in {dt:.2f}s on {cores} threads; GPU vs FP64 gradient "
        f"(rel norm) n {err_n:.2e}, z {err_z:.2e}")
    return {
        "value": geom.N * k_steps / dt,
        "unit": "cell-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 reverse-mode gradient (oracle, torch autodiff) of the first {k_steps} steps of the same "
        f"{geom.N}-cell run from the dry start (recorded dt, loss at the 32 cells nearest the inlet every {every} "
        f"steps); the GPU adjoint of the same slice differs from it by {err_n:.2e} (n) and {err_z:.2e} (z) in "
        f"relative norm",
        "gpu_vs_fp64": {"n": err_n, "z": err_z},
    }


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
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"{k} Heun steps of the same {n}-cell mesh from the GPU state after the timed run "
        f"(FP32 C++/OpenMP oracle; bitwise equal to the GPU on those steps)",
        "fp64_oracle_c1": fp64_c1_rate(),
    }


def cpu_model() -> str:
    """The host CPU's model name."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def fp64_c1_rate(budget_s: float = 5.0) -> dict:
    """The FP64 NumPy oracle on the 4096-cell lake at rest, one process, for about budget_s seconds."""
    from inundation_amd.hip.solver import device_mesh_arrays
    from inundation_amd.workloads import lake_at_rest
    from oracle import ref64

    wl = lake_at_rest()
    m64 = ref64.Mesh64.from_device_arrays(device_mesh_arrays(wl.geom))
    n = wl.geom.N
    s = ref64.State64(wl.h0.astype(np.float64), np.zeros(n), np.zeros(n))
    nm = wl.manning.astype(np.float64)
    k = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        s, _ = ref64.heun_step(m64, s, nm, 0.05, 9.81, wl.dry_tol)
        k += 1
    dt = time.perf_counter() - t0
    log(f"fp64 oracle on the lake at rest: {k} steps in {dt:.2f}s, one process")
    return {"value": k * n / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{k} FP64 NumPy Heun steps (dt 0.05 s) of the {n}-cell lake at rest, one process"}


if __name__ == "__main__":
    sys.exit(main())
