"""ctypes driver of the FP32 oracle ``libcpu_ref.so`` (TEST INFRASTRUCTURE ONLY).

``CpuRef`` takes the same calls as the backend's ``inundation_amd.hip.lib.Context``
(set_params, set_state, set_sources, ctrl_write/read, prologue, retarget,
steps, download, probe), so a test drives both with one script and compares
the results bit for bit.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Any

import numpy as np
from numpy.typing import NDArray

from inundation_amd.tuning import ACTIVE_BLOCK_CELLS

from .blocks_ref import block_adjacency, block_two_ring, edge_workgroup_block_ranges

LIB_PATH = Path(__file__).resolve().parent / "libcpu_ref.so"
_lib: C.CDLL | None = None
_P = C.c_void_p
_U = C.c_uint32
_F = C.c_float


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        from .build import build

        build()
    lib = C.CDLL(str(LIB_PATH))
    sig = {
        "cref_create": (_P, [_U, _U] + [_P] * 12),
        "cref_destroy": (None, [_P]),
        "cref_threads": (None, [C.c_int]),
        "cref_max_threads": (C.c_int, []),
        "cref_set_params": (None, [_P, _F, _F, _F, _U]),
        "cref_set_state": (None, [_P, _P, _P, _P, _P]),
        "cref_set_manning": (None, [_P, _P]),
        "cref_set_sources": (None, [_P, _U, _P, _P]),
        "cref_ctrl_write": (None, [_P, _P]),
        "cref_ctrl_read": (None, [_P, _P]),
        "cref_dt_words": (None, [_P, _P]),
        "cref_prologue": (None, [_P]),
        "cref_retarget": (None, [_P]),
        "cref_steps": (None, [_P, _U]),
        "cref_download": (None, [_P, C.c_int, _P]),
        "cref_probe": (None, [_U, _P, _P, _F, _F]),
        "cref_probe_open": (None, [_U, _P, _P, _P, _P, _F, _F]),
        "cref_set_open_stretches": (None, [_P, _U, _P]),
        "cref_set_open_conditions": (None, [_P, _P, _P]),
        "cref_open_count": (_U, [_P]),
        "cref_open_tally": (None, [_P, _P]),
        "cref_probe_clock": (None, [C.c_int, _P, _P]),
        "cref_upload": (None, [_P, C.c_int, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a: NDArray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def set_threads(n: int) -> None:
    load().cref_threads(int(n))


def max_threads() -> int:
    return int(load().cref_max_threads())


def probe(pairs: NDArray[np.float32], g: float, dry_tol: float) -> NDArray[np.float32]:
    p = np.ascontiguousarray(pairs, dtype=np.float32).reshape(-1, 2)
    out = np.empty((p.shape[0], 7), dtype=np.float32)
    load().cref_probe(p.shape[0], _p(p), _p(out), g, dry_tol)
    return out


def probe_open(cases, kinds, values, g: float, dry_tol: float) -> NDArray[np.float32]:  # noqa: ANN001
    """Cases (h, hu, hv, nx, ny, fh) -> ghost (h, hu, hv) and boundary mass flux."""
    c = np.ascontiguousarray(cases, dtype=np.float32).reshape(-1, 6)
    k = np.ascontiguousarray(kinds, dtype=np.uint32).ravel()
    v = np.ascontiguousarray(values, dtype=np.float32).ravel()
    assert k.size == v.size == c.shape[0]
    out = np.empty((c.shape[0], 4), dtype=np.float32)
    load().cref_probe_open(c.shape[0], _p(c), _p(k), _p(v), _p(out), g, dry_tol)
    return out


def probe_clock(mode: int, ctrl: Any, dt_words: NDArray[np.uint32]) -> tuple[Any, NDArray[np.uint32]]:
    """step_control (mode 0 START, 1 STEP, 2 RETARGET) on a copy of ``ctrl`` (an ``InuCtrl``) and the
    dt words (prologue bound bits, next-step bound bits, 0, 0); returns the updated copies."""
    out = type(ctrl).from_buffer_copy(ctrl)
    dt = np.zeros(4, dtype=np.uint32)
    dt[: len(dt_words)] = np.asarray(dt_words, dtype=np.uint32)
    load().cref_probe_clock(int(mode), C.addressof(out), _p(dt))
    return out, dt


class CpuRef:
    """One oracle instance over one mesh, given in the device edge order."""

    def __init__(self, m: dict[str, Any]) -> None:
        """``m`` as returned by ``inundation_amd.hip.device_mesh_arrays(geom)``."""
        self._lib = load()
        n = int(m["area"].size)
        e = int(m["cell_l"].size)
        self.N, self.E = n, e
        self.NB = -(-n // ACTIVE_BLOCK_CELLS)
        erange = edge_workgroup_block_ranges(m["cell_l"], 256, ACTIVE_BLOCK_CELLS)
        bptr, bidx = block_two_ring(*block_adjacency(m["cell_l"], m["cell_r"], n, ACTIVE_BLOCK_CELLS))
        self._keep = [
            np.ascontiguousarray(x, dtype=dt)
            for x, dt in (
                (m["cell_l"], np.int32),
                (m["cell_r"], np.int32),
                (m["nx"], np.float32),
                (m["ny"], np.float32),
                (m["len"], np.float32),
                (m["dz2"], np.float32),
                (m["area"], np.float32),
                (m["cptr"], np.uint32),
                (m["cidx"], np.uint32),
                (erange.ravel(), np.uint32),
                (bptr, np.uint32),
                (bidx if bidx.size else np.zeros(1, np.uint32), np.uint32),
            )
        ]
        self._h = self._lib.cref_create(n, e, *[_p(x) for x in self._keep])

    def __del__(self) -> None:  # pragma: no cover
        if getattr(self, "_h", None):
            self._lib.cref_destroy(self._h)
            self._h = None

    def set_params(self, g: float, dry_tol: float, cfl: float, flags: int) -> None:
        self._lib.cref_set_params(self._h, g, dry_tol, cfl, flags)

    def set_state(self, h, hu, hv, h_lo=None) -> None:  # noqa: ANN001
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (h, hu, hv)]
        lo = None if h_lo is None else np.ascontiguousarray(h_lo, dtype=np.float32)
        self._lib.cref_set_state(self._h, *[_p(x) for x in a], None if lo is None else _p(lo))

    def set_manning(self, n) -> None:  # noqa: ANN001
        a = np.ascontiguousarray(n, dtype=np.float32)
        self._lib.cref_set_manning(self._h, _p(a))

    def set_sources(self, cells, rates) -> None:  # noqa: ANN001
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        r = np.ascontiguousarray(rates, dtype=np.float32)
        keep = r != 0.0
        c, r = np.ascontiguousarray(c[keep]), np.ascontiguousarray(r[keep])
        order = np.argsort(c, kind="stable")
        c, r = np.ascontiguousarray(c[order]), np.ascontiguousarray(r[order])
        self._src = (c, r)
        self._lib.cref_set_sources(self._h, c.size, _p(c) if c.size else None, _p(r) if r.size else None)

    def ctrl_write(self, ctrl: Any) -> None:
        buf = (C.c_uint8 * 48).from_buffer_copy(bytes(ctrl))
        self._lib.cref_ctrl_write(self._h, C.addressof(buf))

    def ctrl_read(self) -> Any:
        from inundation_amd.hip.lib import InuCtrl

        c = InuCtrl()
        self._lib.cref_ctrl_read(self._h, C.addressof(c))
        return c

    def dt_words(self) -> NDArray[np.uint32]:
        out = np.zeros(4, dtype=np.uint32)
        self._lib.cref_dt_words(self._h, _p(out))
        return out

    def prologue(self) -> None:
        self._lib.cref_prologue(self._h)

    def retarget(self) -> None:
        self._lib.cref_retarget(self._h)

    def steps(self, n: int, use_graph: bool = True) -> None:
        del use_graph
        self._lib.cref_steps(self._h, n)

    @property
    def n_open(self) -> int:
        return int(self._lib.cref_open_count(self._h))

    def set_open_stretches(self, stretch_ptr) -> None:  # noqa: ANN001
        p = np.ascontiguousarray(stretch_ptr, dtype=np.uint32)
        self._n_stretch = p.size - 1
        self._lib.cref_set_open_stretches(self._h, p.size - 1, _p(p))

    def set_open_conditions(self, kinds, values) -> None:  # noqa: ANN001
        k = np.ascontiguousarray(kinds, dtype=np.uint32)
        v = np.ascontiguousarray(values, dtype=np.float32)
        assert k.size == v.size == self.n_open
        self._lib.cref_set_open_conditions(self._h, _p(k), _p(v))

    def open_tally(self) -> NDArray[np.float32]:
        """Per stretch: out (hi, lo), in (hi, lo) as float32 words."""
        n = getattr(self, "_n_stretch", 1 if self.n_open else 0)
        out = np.zeros(4 * n, dtype=np.float32)
        if n:
            self._lib.cref_open_tally(self._h, _p(out))
        return out.reshape(n, 4)

    def open_volumes(self, n_stretches: int | None = None) -> NDArray[np.float64]:
        """Per stretch (out, in) [m^3], each as hi + lo in FP64."""
        t = self.open_tally().astype(np.float64)
        out = np.stack([t[:, 0] + t[:, 1], t[:, 2] + t[:, 3]], axis=1)
        assert n_stretches is None or n_stretches == out.shape[0]
        return out

    def download(self, which: int, n: int, dtype: type = np.float32) -> NDArray:
        out = np.empty(n, dtype=dtype)
        self._lib.cref_download(self._h, which, _p(out))
        return out

    def upload(self, which: int, values) -> None:  # noqa: ANN001
        a = np.ascontiguousarray(values, dtype=np.float32)
        assert a.size == self.N and 0 <= which <= 8
        self._lib.cref_upload(self._h, which, _p(a))

    def probe(self, pairs, g: float, dry_tol: float) -> NDArray[np.float32]:  # noqa: ANN001
        return probe(pairs, g, dry_tol)

    def close(self) -> None:
        """Nothing to release early (the library handle goes with the object)."""


__all__ = ["CpuRef", "load", "max_threads", "probe", "probe_open", "set_threads"]
