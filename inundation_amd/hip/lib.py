"""ctypes binding of ``libinundation_hip.so`` (declared in ``include/inundation_hip.h``).

The library is loaded from the package directory only; a missing library is a
hard error (there is no CPU fallback on the product path).
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Sequence
from pathlib import Path
from typing import Any

import numpy as np
from numpy.typing import NDArray

LIB_PATH = Path(__file__).resolve().parent.parent / "libinundation_hip.so"

OK = 0
EINVAL = -1
EHIP = -2
ESTATE = -3
ERANGE = -5
ENODEV = -6

FLAG_DETERMINISTIC = 1
FLAG_COMPENSATED = 2
FLAG_ACTIVE_SET = 4
FLAG_TRACK_CLAMP = 8
FLAG_EXACT_ORDER = 16

OPEN_WALL, OPEN_FREE, OPEN_NORMAL, OPEN_STAGE, OPEN_INFLOW = 0, 1, 2, 3, 4
FIELD_H = 0
FIELD_HU = 1
FIELD_HV = 2
FIELD_H_LO = 3
FIELD_H1 = 4
FIELD_HU1 = 5
FIELD_HV1 = 6
FIELD_H1_LO = 7
FIELD_CLAMP = 8
FIELD_BLOCK_WET = 9
FIELD_BLOCK_ACTIVE = 10
FIELD_BLOCK_FORCED = 11

# Every entry point the header declares, for the export check.
EXPORTS = (
    "inu_version",
    "inu_create",
    "inu_destroy",
    "inu_last_error",
    "inu_set_mesh",
    "inu_set_batch",
    "inu_scenario_count",
    "inu_scenario_info",
    "inu_mesh_info",
    "inu_set_params",
    "inu_set_state",
    "inu_set_manning",
    "inu_set_sources",
    "inu_set_source_rates",
    "inu_ctrl_write",
    "inu_ctrl_read",
    "inu_dt_words",
    "inu_activity",
    "inu_prologue",
    "inu_retarget",
    "inu_steps",
    "inu_steps_timed",
    "inu_download",
    "inu_upload",
    "inu_probe",
    "inu_stream_copy",
    "inu_autotune",
    "inu_set_tile",
    "inu_set_bed",
    "inu_adj_begin",
    "inu_adj_steps",
    "inu_adj_series",
    "inu_adj_backward",
    "inu_adj_replay",
    "inu_adj_dt_log",
    "inu_adj_stats",
    "inu_adj_set_thin",
    "inu_adj_set_slots",
    "inu_adj_replayed",
    "inu_adj_schedule_cost",
    "inu_set_open_stretches",
    "inu_set_open_conditions",
    "inu_open_volumes",
    "inu_probe_open",
    "inu_adj_end",
    "inu_probe_vjp",
    "inu_probe_clock",
    "inu_probe_rcp",
    "inu_stream_bw",
    "inu_set_source_rates_noalloc",
)


class InuCtrl(C.Structure):
    """``inu_ctrl``: the 48-byte resident control block (gpu_resident_batch.py:65-77)."""

    _fields_ = [
        ("now_ticks", C.c_uint64),
        ("stop_ticks", C.c_uint64),
        ("steps", C.c_uint64),
        ("dt_ticks", C.c_uint32),
        ("dt_max_ticks", C.c_uint32),
        ("dry_dt_ticks", C.c_uint32),
        ("last_raw", C.c_float),
        ("cfl_safety", C.c_float),
        ("pad", C.c_uint32),
    ]


class InuMesh(C.Structure):
    """``inu_mesh``: one mesh of a batch, as the arguments of ``inu_set_mesh``."""

    _fields_ = [
        ("n_cells", C.c_uint32),
        ("n_edges", C.c_uint32),
        ("edge_cell_l", C.c_void_p),
        ("edge_cell_r", C.c_void_p),
        ("nx", C.c_void_p),
        ("ny", C.c_void_p),
        ("len", C.c_void_p),
        ("edge_dz", C.c_void_p),
        ("cell_area", C.c_void_p),
        ("cell_edge_ptr", C.c_void_p),
        ("cell_edge_idx", C.c_void_p),
    ]


class HipError(RuntimeError):
    """A HIP runtime failure inside the library."""


_P = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int
_F = C.c_float
_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the in-tree library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m inundation_amd.build` "
            "(the HIP backend has no CPU fallback)"
        )
    lib = C.CDLL(str(LIB_PATH))
    sig = {
        "inu_version": (_I, []),
        "inu_create": (_I, [_I, C.POINTER(_P)]),
        "inu_destroy": (None, [_P]),
        "inu_last_error": (C.c_char_p, [_P]),
        "inu_set_mesh": (_I, [_P, _U32, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "inu_set_batch": (_I, [_P, _U32, _P, _U32, _P]),
        "inu_scenario_count": (_I, [_P]),
        "inu_scenario_info": (_I, [_P, _U32, _P]),
        "inu_mesh_info": (_I, [_P, _P]),
        "inu_set_params": (_I, [_P, _F, _F, _F, _U32]),
        "inu_set_state": (_I, [_P, _P, _P, _P, _P]),
        "inu_set_manning": (_I, [_P, _P]),
        "inu_set_sources": (_I, [_P, _U32, _P, _P]),
        "inu_set_source_rates": (_I, [_P, _U32, _P, _P]),
        "inu_ctrl_write": (_I, [_P, C.POINTER(InuCtrl)]),
        "inu_ctrl_read": (_I, [_P, C.POINTER(InuCtrl)]),
        "inu_dt_words": (_I, [_P, _P]),
        "inu_activity": (_I, [_P, _P, _I]),
        "inu_prologue": (_I, [_P]),
        "inu_retarget": (_I, [_P]),
        "inu_steps": (_I, [_P, _U32, _I]),
        "inu_steps_timed": (_I, [_P, _U32, _P]),
        "inu_download": (_I, [_P, _I, _P]),
        "inu_upload": (_I, [_P, _I, _P]),
        "inu_probe": (_I, [_P, _U32, _P, _P, _F, _F]),
        "inu_stream_copy": (_I, [_P, C.c_uint64, _U32, _P]),
        "inu_autotune": (_I, [_P, _U32, _P]),
        "inu_set_tile": (_I, [_P, _U32]),
        "inu_set_bed": (_I, [_P, _P]),
        "inu_adj_begin": (_I, [_P, _U32, _U32, _U32, _P, _U32]),
        "inu_adj_steps": (_I, [_P, _U32]),
        "inu_adj_series": (_I, [_P, _P]),
        "inu_adj_backward": (_I, [_P, _P, _P, _P, _P]),
        "inu_adj_replay": (_I, [_P, _P]),
        "inu_adj_dt_log": (_I, [_P, _P]),
        "inu_adj_stats": (_I, [_P, _P]),
        "inu_adj_set_thin": (_I, [_P, _F]),
        "inu_adj_set_slots": (_I, [_P, _U32]),
        "inu_adj_replayed": (_I, [_P, _P]),
        "inu_adj_schedule_cost": (_I, [_U32, _U32, _P]),
        "inu_set_open_stretches": (_I, [_P, _U32, _P]),
        "inu_set_open_conditions": (_I, [_P, _U32, _P, _P]),
        "inu_open_volumes": (_I, [_P, _U32, _P]),
        "inu_probe_open": (_I, [_P, _U32, _P, _P, _P, _P, _F, _F]),
        "inu_adj_end": (_I, [_P]),
        "inu_probe_vjp": (_I, [_P, _U32, _P, _P, _F, _F]),
        "inu_probe_clock": (_I, [_P, _I, C.POINTER(InuCtrl), _P]),
        "inu_probe_rcp": (_I, [_P, _U32, _U32, _P, _P]),
        "inu_stream_bw": (_I, [_P, C.c_uint64, _U32, _I, _U32, _U32, _P]),
        "inu_set_source_rates_noalloc": (_I, [_P, _U32, _P, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def schedule_cost(n_segments: int, slots: int) -> int:
    """Segments the binomial checkpoint schedule advances without tape (host only; see inu_adj_schedule_cost)."""
    out = C.c_ulonglong(0)
    rc = load().inu_adj_schedule_cost(int(n_segments), int(slots), C.byref(out))
    if rc != OK:
        raise ValueError(f"schedule_cost: segments in [1, 2048] and slots in [1, 64] (code {rc})")
    return int(out.value)


def ptr(a: NDArray) -> int:
    """Address of a C-contiguous array's data."""
    if not a.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


class Context:
    """Owns one ``inu_ctx``: one device copy of a mesh, its state and stream."""

    def __init__(self, device: int = 0) -> None:
        self._lib = load()
        h = _P()
        rc = self._lib.inu_create(device, C.byref(h))
        if rc != OK:
            if rc == ENODEV:
                raise RuntimeError("no HIP device available for the hip_resident backend")
            raise RuntimeError(f"inu_create failed with code {rc}")
        self._h = h
        self.n_scenarios = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.inu_destroy(self._h)
            self._h = None

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int) -> None:
        """Map a negative return code to ValueError (bad argument) or RuntimeError."""
        if rc == OK:
            return
        msg = self._lib.inu_last_error(self._h).decode(errors="replace")
        if rc == EINVAL:
            raise ValueError(msg)
        if rc == EHIP:
            raise HipError(msg)
        raise RuntimeError(f"{msg} (code {rc})")

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def call(self, name: str, *args: object) -> None:
        self.check(getattr(self._lib, name)(self._h, *args))

    @staticmethod
    def _mesh_arrays(m: dict[str, Any]) -> list[NDArray]:
        return [
            np.ascontiguousarray(m["cell_l"], dtype=np.int32),
            np.ascontiguousarray(m["cell_r"], dtype=np.int32),
            np.ascontiguousarray(m["nx"], dtype=np.float32),
            np.ascontiguousarray(m["ny"], dtype=np.float32),
            np.ascontiguousarray(m["len"], dtype=np.float32),
            np.ascontiguousarray(m["dz2"], dtype=np.float32),
            np.ascontiguousarray(m["area"], dtype=np.float32),
            np.ascontiguousarray(m["cptr"], dtype=np.uint32),
            np.ascontiguousarray(m["cidx"], dtype=np.uint32),
        ]

    def set_mesh(
        self,
        cell_l: NDArray[np.int32],
        cell_r: NDArray[np.int32],
        nx: NDArray[np.float32],
        ny: NDArray[np.float32],
        length: NDArray[np.float32],
        dz2: NDArray[np.float32],
        area: NDArray[np.float32],
        cptr: NDArray[np.uint32],
        cidx: NDArray[np.uint32],
    ) -> None:
        arrs = self._mesh_arrays(
            {"cell_l": cell_l, "cell_r": cell_r, "nx": nx, "ny": ny, "len": length, "dz2": dz2, "area": area,
             "cptr": cptr, "cidx": cidx}
        )
        self.call("inu_set_mesh", arrs[6].size, arrs[0].size, *[ptr(a) for a in arrs])
        self.n_scenarios = 1

    def set_batch(self, meshes: Sequence[dict[str, Any]], scenario_mesh: Sequence[int]) -> None:
        """Scenario s on ``meshes[scenario_mesh[s]]`` (each as ``device_mesh_arrays`` returns it)."""
        keep = [self._mesh_arrays(m) for m in meshes]
        descs = (InuMesh * len(keep))()
        for d, arrs in zip(descs, keep):
            d.n_cells, d.n_edges = arrs[6].size, arrs[0].size
            (d.edge_cell_l, d.edge_cell_r, d.nx, d.ny, d.len, d.edge_dz, d.cell_area, d.cell_edge_ptr,
             d.cell_edge_idx) = [ptr(a) for a in arrs]
        sm = np.ascontiguousarray(scenario_mesh, dtype=np.uint32)
        self.call("inu_set_batch", len(keep), C.addressof(descs), sm.size, ptr(sm))
        self.n_scenarios = int(sm.size)

    def scenario_info(self, s: int) -> dict[str, int]:
        out = np.zeros(4, dtype=np.uint32)
        self.call("inu_scenario_info", s, ptr(out))
        return {"cell0": int(out[0]), "cells": int(out[1]), "block0": int(out[2]), "blocks": int(out[3])}

    def mesh_info(self) -> dict[str, int]:
        out = np.zeros(6, dtype=np.uint32)
        self.call("inu_mesh_info", ptr(out))
        keys = ("N", "E", "S", "NB", "max_local", "grid_stage")
        return {k: int(v) for k, v in zip(keys, out.tolist())}

    def set_params(self, g: float, dry_tol: float, cfl: float, flags: int) -> None:
        self.call("inu_set_params", g, dry_tol, cfl, flags)

    def set_state(
        self,
        h: NDArray[np.float32],
        hu: NDArray[np.float32],
        hv: NDArray[np.float32],
        h_lo: NDArray[np.float32] | None = None,
    ) -> None:
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (h, hu, hv)]
        lo = None if h_lo is None else np.ascontiguousarray(h_lo, dtype=np.float32)
        self.call("inu_set_state", *[ptr(x) for x in a], None if lo is None else ptr(lo))

    def set_manning(self, n: NDArray[np.float32]) -> None:
        a = np.ascontiguousarray(n, dtype=np.float32)
        self.call("inu_set_manning", ptr(a))

    def set_sources(self, cells: NDArray[np.uint32], rates: NDArray[np.float32]) -> None:
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        r = np.ascontiguousarray(rates, dtype=np.float32)
        if c.size != r.size:
            raise ValueError("cells and rates differ in length")
        self.call("inu_set_sources", c.size, ptr(c) if c.size else None, ptr(r) if r.size else None)

    def set_source_rates(self, cells: NDArray[np.uint32], rates: NDArray[np.float32]) -> None:
        """New rates for cells already in the source set (zero turns a source off)."""
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        r = np.ascontiguousarray(rates, dtype=np.float32)
        if c.size != r.size:
            raise ValueError("cells and rates differ in length")
        self.call("inu_set_source_rates", c.size, ptr(c) if c.size else None, ptr(r) if r.size else None)

    def set_source_rates_noalloc(self, cells: NDArray[np.uint32], rates: NDArray[np.float32]) -> None:
        """``set_source_rates`` with no device allocation (at most 65536 rates; see the header)."""
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        r = np.ascontiguousarray(rates, dtype=np.float32)
        if c.size != r.size:
            raise ValueError("cells and rates differ in length")
        self.call("inu_set_source_rates_noalloc", c.size, ptr(c) if c.size else None, ptr(r) if r.size else None)

    def ctrl_write(self, ctrl: InuCtrl | Sequence[InuCtrl]) -> None:
        """One control block, or one per scenario."""
        items = [ctrl] if isinstance(ctrl, InuCtrl) else list(ctrl)
        if len(items) != self.n_scenarios:
            raise ValueError(f"{len(items)} control blocks for {self.n_scenarios} scenarios")
        arr = (InuCtrl * len(items))(*items)
        self.call("inu_ctrl_write", arr)

    def ctrl_read(self) -> InuCtrl:
        """The control block of a one-scenario context."""
        if self.n_scenarios != 1:
            raise ValueError("ctrl_read() is for one scenario; use ctrls_read()")
        return self.ctrls_read()[0]

    def ctrls_read(self) -> list[InuCtrl]:
        arr = (InuCtrl * self.n_scenarios)()
        self.call("inu_ctrl_read", arr)
        return list(arr)

    def dt_words(self) -> NDArray[np.uint32]:
        """Four words per scenario: this step's dt bits, next-step bound bits, 0, 0 (flat when one scenario)."""
        out = np.zeros(4 * self.n_scenarios, dtype=np.uint32)
        self.call("inu_dt_words", ptr(out))
        return out if self.n_scenarios == 1 else out.reshape(self.n_scenarios, 4)

    def activity(self, reset: bool = False) -> dict[str, int]:
        """Active cells, owned edges and slots summed over the real steps, and the step count."""
        out = np.zeros(4, dtype=np.uint64)
        self.call("inu_activity", ptr(out), 1 if reset else 0)
        return {"cells": int(out[0]), "edges": int(out[1]), "slots": int(out[2]), "steps": int(out[3])}

    def prologue(self) -> None:
        self.call("inu_prologue")

    def retarget(self) -> None:
        self.call("inu_retarget")

    def steps(self, n: int, use_graph: bool = True) -> None:
        self.call("inu_steps", n, 1 if use_graph else 0)

    def steps_timed(self, n: int) -> tuple[float, float, float]:
        out = np.zeros(3, dtype=np.float32)
        self.call("inu_steps_timed", n, ptr(out))
        return float(out[0]), float(out[1]), float(out[2])

    def download(self, which: int, n: int, dtype: type = np.float32) -> NDArray:
        out = np.empty(n, dtype=dtype)
        self.call("inu_download", which, ptr(out))
        return out

    def upload(self, which: int, values: NDArray[np.float32]) -> None:
        a = np.ascontiguousarray(values, dtype=np.float32)
        self.call("inu_upload", which, ptr(a))

    def stream_copy(self, n_bytes: int, reps: int = 5) -> float:
        """Mean milliseconds of one plain device copy of ``n_bytes`` (read and written)."""
        ms = np.zeros(1, dtype=np.float32)
        self.call("inu_stream_copy", int(n_bytes), int(reps), ptr(ms))
        return float(ms[0])

    def stream_bw(self, n_bytes: int, kind: int, unroll: int = 4, wg_per_cu: int = 8, reps: int = 5) -> float:
        """Mean milliseconds of one bandwidth probe (kind 0 copy, 1 read, 2 write, 3 hipMemcpy) of ``n_bytes``."""
        ms = np.zeros(1, dtype=np.float32)
        self.call("inu_stream_bw", int(n_bytes), int(reps), int(kind), int(unroll), int(wg_per_cu), ptr(ms))
        return float(ms[0])

    def autotune(self, reps: int = 5) -> dict[str, Any]:
        """Pick the stage kernels' tile level by timing the predictor; returns the choice and the timings."""
        out = np.zeros(4, dtype=np.uint32)
        self.call("inu_autotune", int(reps), ptr(out))
        return {"tile": int(out[0]), "us_per_launch": [int(x) for x in out[1:]]}

    def set_tile(self, tile: int) -> None:
        self.call("inu_set_tile", int(tile))

    def set_bed(self, z: NDArray[np.float32]) -> None:
        """Bed elevation per cell; every edge's bed step is recomputed."""
        a = np.ascontiguousarray(z, dtype=np.float32)
        self.call("inu_set_bed", ptr(a))

    def adj_begin(self, max_steps: int, seg: int, gauges: NDArray[np.uint32], every: int) -> None:
        g = np.ascontiguousarray(gauges, dtype=np.uint32)
        self._adj = (int(max_steps) // int(every), g.size)
        self.call("inu_adj_begin", int(max_steps), int(seg), g.size, ptr(g), int(every))

    def adj_steps(self, n: int) -> None:
        self.call("inu_adj_steps", int(n))

    def adj_series(self) -> NDArray[np.float32]:
        out = np.zeros(self._adj, dtype=np.float32)
        self.call("inu_adj_series", ptr(out))
        return out

    def adj_backward(self, seeds: NDArray[np.float32], n_cells: int) -> tuple[NDArray, NDArray, NDArray]:
        """(d loss / d n, d loss / d z, d loss / d initial (h, hu, hv)) for seeds = d loss / d series."""
        sd = np.ascontiguousarray(seeds, dtype=np.float32).reshape(self._adj)
        gn = np.zeros(n_cells, np.float32)
        gz = np.zeros(n_cells, np.float32)
        gs = np.zeros((3, n_cells), np.float32)
        self.call("inu_adj_backward", ptr(sd), ptr(gn), ptr(gz), ptr(gs))
        return gn, gz, gs

    def adj_replay(self) -> NDArray[np.float32]:
        """Gauge series of the recorded steps replayed with the current n and bed."""
        out = np.zeros(self._adj, dtype=np.float32)
        self.call("inu_adj_replay", ptr(out))
        return out

    def adj_dts(self, max_steps: int) -> NDArray[np.float32]:
        """Recorded dt of each step (0 past the last recorded step)."""
        out = np.zeros(int(max_steps), dtype=np.uint32)
        self.call("inu_adj_dt_log", ptr(out))
        return out.view(np.float32)

    def adj_stats(self) -> dict[str, int]:
        """Counters of the last backward pass (see inu_adj_stats)."""
        out = np.zeros(3, dtype=np.uint64)
        self.call("inu_adj_stats", ptr(out))
        return {"listed_block_steps": int(out[0]), "idle_block_steps": int(out[1]), "boundary_film_edges": int(out[2])}

    def adj_set_thin(self, depth: float) -> None:
        """Face depth below which the reverse pass holds a side's velocity and celerity."""
        self.call("inu_adj_set_thin", float(depth))

    def adj_set_slots(self, slots: int) -> None:
        """Checkpoint slots of the next recording: 0 uniform (every seg steps), else binomial with that many."""
        self.call("inu_adj_set_slots", int(slots))

    def adj_replayed(self) -> dict[str, int]:
        """Steps the last backward pass replayed without tape and with tape."""
        out = np.zeros(2, dtype=np.uint64)
        self.call("inu_adj_replayed", ptr(out))
        return {"advanced": int(out[0]), "taped": int(out[1])}

    def adj_end(self) -> None:
        self.call("inu_adj_end")

    def set_open_stretches(self, stretch_ptr: NDArray[np.uint32]) -> None:
        """Open edges [ptr[s], ptr[s + 1]) form stretch s."""
        p = np.ascontiguousarray(stretch_ptr, dtype=np.uint32)
        self.call("inu_set_open_stretches", p.size - 1, ptr(p))

    def set_open_conditions(self, kinds: NDArray[np.uint32], values: NDArray[np.float32]) -> None:
        k = np.ascontiguousarray(kinds, dtype=np.uint32)
        v = np.ascontiguousarray(values, dtype=np.float32)
        if k.size != v.size:
            raise ValueError("one value per open edge")
        self.call("inu_set_open_conditions", k.size, ptr(k) if k.size else None, ptr(v) if v.size else None)

    def open_volumes(self, n_stretches: int) -> NDArray[np.float64]:
        """(stretches, 2) volume [m^3] out of and into the mesh through each stretch."""
        out = np.zeros((int(n_stretches), 2), dtype=np.float64)
        self.call("inu_open_volumes", out.shape[0], ptr(out) if out.size else None)
        return out

    def probe_vjp(self, rows: NDArray[np.float32], g: float, dry_tol: float) -> NDArray[np.float32]:
        r = np.ascontiguousarray(rows, dtype=np.float32).reshape(-1, 17)
        out = np.empty((r.shape[0], 8), dtype=np.float32)
        self.call("inu_probe_vjp", r.shape[0], ptr(r), ptr(out), g, dry_tol)
        return out

    def probe_open(
        self, cases: NDArray[np.float32], kinds: NDArray[np.uint32], values: NDArray[np.float32], g: float,
        dry_tol: float,
    ) -> NDArray[np.float32]:
        """Cases (h, hu, hv, nx, ny, fh) -> ghost (h, hu, hv) and boundary mass flux."""
        c = np.ascontiguousarray(cases, dtype=np.float32).reshape(-1, 6)
        k = np.ascontiguousarray(kinds, dtype=np.uint32).ravel()
        v = np.ascontiguousarray(values, dtype=np.float32).ravel()
        if not (k.size == v.size == c.shape[0]):
            raise ValueError("one kind and one value per case")
        out = np.empty((c.shape[0], 4), dtype=np.float32)
        self.call("inu_probe_open", c.shape[0], ptr(c), ptr(k), ptr(v), ptr(out), g, dry_tol)
        return out

    def probe_clock(self, mode: int, ctrl: InuCtrl, dt_words: Sequence[int]) -> tuple[InuCtrl, NDArray[np.uint32]]:
        """step_control (0 START, 1 STEP, 2 RETARGET) on a copy of ``ctrl`` and the dt words (prologue
        bound bits, next-step bound bits); returns the updated copies. The context's clock is untouched."""
        out = InuCtrl.from_buffer_copy(ctrl)
        dt = np.zeros(4, dtype=np.uint32)
        dt[: len(dt_words)] = np.asarray(dt_words, dtype=np.uint32)
        self.call("inu_probe_clock", int(mode), C.byref(out), ptr(dt))
        return out, dt

    def probe_rcp(self, lo: int, hi: int) -> tuple[int, int]:
        """Bit patterns in [lo, hi) and their negatives where the fast reciprocal differs from the IEEE
        division: (count, smallest such pattern or 0xffffffff)."""
        cnt = np.zeros(1, dtype=np.uint64)
        first = np.zeros(1, dtype=np.uint32)
        self.call("inu_probe_rcp", int(lo), int(hi), ptr(cnt), ptr(first))
        return int(cnt[0]), int(first[0])

    def probe(self, pairs: NDArray[np.float32], g: float, dry_tol: float) -> NDArray[np.float32]:
        p = np.ascontiguousarray(pairs, dtype=np.float32).reshape(-1, 2)
        out = np.empty((p.shape[0], 7), dtype=np.float32)
        self.call("inu_probe", p.shape[0], ptr(p), ptr(out), g, dry_tol)
        return out
