"""Physical defaults, the integer simulated clock and the active-set block size.

These values are part of the step contract: the kernels, the FP32 and FP64
oracles and the workflow all read them from here, so they cannot drift apart.

Reference: ``inundation/tuning.py:26-69`` (constants) and ``:98-152`` (helpers).
"""

from __future__ import annotations

import math
import struct
from typing import TYPE_CHECKING

if TYPE_CHECKING:
    from collections.abc import Iterator

GRAVITY_G = 9.81
DRY_TOL_DEFAULT = 1e-4
# dt_i = CFL * A_i / sum_j(max|S_j| * len_j) over the cell's edges, with the wave
# speeds the flux assumed (reference tuning.py:28-33).
CFL_DEFAULT = 0.8
# CFL number of the host-driven backends, which reuse one dt for cfl_interval
# steps (reference tuning.py:34-36); their host loops stop below
# CFL_EPSILON_MIN and ignore bounds above CFL_SANITY_MAX (an all-dry mesh)
# (tuning.py:43-45).
CFL_HOST_DRIVEN = 0.45
CFL_EPSILON_MIN = 1e-10
CFL_SANITY_MAX = 1e10
# Below FILM_DEPTHS * dry_tol velocities are desingularised and film momentum
# decays at a rate (reference tuning.py:37-39).
FILM_DEPTHS = 4.0

# A mesh polygon needs at least this many vertices (reference tuning.py:23).
MIN_POLYGON_VERTICES = 3

# Cells below this area are refused by build_geometry; the CFL reductions skip
# cells at or below it (reference tuning.py:88-92).
MIN_CELL_AREA = 1e-10
POLYGON_CENTROID_EPSILON = 1e-20
EDGE_LENGTH_EPSILON = 1e-30

# Simulated time is a uint64 count of 2^-20 s ticks; every dt is a whole number
# of ticks no larger than 2^24, so float(ticks) * 2^-20 is exact in float32
# (reference tuning.py:47-51).
SIM_TICKS_PER_SECOND = 1 << 20
SIM_DT_MAX_TICKS = 1 << 24

# Active-set block: consecutive Hilbert-ordered cells handled by one workgroup
# (reference tuning.py:53-55).
ACTIVE_BLOCK_CELLS = 256

# Next-step CFL slot sentinel (reference solver.py:183).
DT_SENTINEL = 1e30

# Backends whose run() returns hu/hv snapshots (reference tuning.py:57-61 plus
# this backend). The reference's backend names run on the HIP kernels here:
# gpu_resident_batch as hip_resident, baseline and fixed_dt_batch_barrier as
# the host-driven loops.
MOMENTUM_SNAPSHOT_BACKENDS = frozenset({"hip_resident", "gpu_resident_batch", "baseline"})
# Backends that take open boundaries (reference tuning.py:61).
OPEN_BOUNDARY_BACKENDS = frozenset({"hip_resident", "gpu_resident_batch", "baseline"})

# Open boundary conditions, the word per open edge (reference tuning.py:64-68).
OPEN_WALL = 0
OPEN_FREE = 1
OPEN_NORMAL = 2
OPEN_STAGE = 3
OPEN_INFLOW = 4


def uint_as_float_bits(n: int) -> float:
    """A 32-bit count as the float whose bit pattern it is (reference tuning.py:98-108)."""
    if not 0 <= n < 2**32:
        raise ValueError(f"count must fit in 32 bits, got {n}")
    return float(struct.unpack("<f", struct.pack("<I", n))[0])


def f32_bits(value: float) -> int:
    """Bit pattern of ``value`` rounded to float32."""
    return struct.unpack("<I", struct.pack("<f", value))[0]


def bits_f32(word: int) -> float:
    """The float32 whose bit pattern is ``word``."""
    return struct.unpack("<f", struct.pack("<I", word & 0xFFFFFFFF))[0]


def seconds_to_ticks(seconds: float) -> int:
    """Simulated seconds to the nearest whole tick (reference tuning.py:121-125)."""
    if not math.isfinite(seconds) or seconds < 0.0:
        raise ValueError(f"simulated time must be finite and non-negative, got {seconds}")
    return round(seconds * SIM_TICKS_PER_SECOND)


def ticks_to_seconds(ticks: int) -> float:
    """Whole ticks to simulated seconds as float64 (reference tuning.py:128-130)."""
    return ticks / SIM_TICKS_PER_SECOND


def dt_to_ticks(dt: float) -> int:
    """A timestep to whole ticks, rounded down so a bound is never exceeded.

    Reference tuning.py:133-142: 1 <= ticks <= 2^24 or ValueError.
    """
    if not math.isfinite(dt) or dt <= 0.0:
        raise ValueError(f"dt must be finite and positive, got {dt}")
    ticks = math.floor(dt * SIM_TICKS_PER_SECOND)
    if ticks < 1 or ticks > SIM_DT_MAX_TICKS:
        raise ValueError(
            f"dt must be between one tick and {SIM_DT_MAX_TICKS / SIM_TICKS_PER_SECOND} s, got {dt}"
        )
    return ticks


def output_window_stops(start_ticks: int, end_ticks: int, interval_ticks: int) -> Iterator[int]:
    """Stop tick of each output window: every interval after the start, then the end.

    Reference tuning.py:145-152.
    """
    if interval_ticks <= 0:
        raise ValueError(f"output interval must be positive, got {interval_ticks} ticks")
    stop = start_ticks
    while stop < end_ticks:
        stop = min(stop + interval_ticks, end_ticks)
        yield stop


__all__ = [
    "ACTIVE_BLOCK_CELLS",
    "CFL_DEFAULT",
    "CFL_EPSILON_MIN",
    "CFL_HOST_DRIVEN",
    "CFL_SANITY_MAX",
    "DRY_TOL_DEFAULT",
    "DT_SENTINEL",
    "FILM_DEPTHS",
    "GRAVITY_G",
    "MIN_CELL_AREA",
    "SIM_DT_MAX_TICKS",
    "SIM_TICKS_PER_SECOND",
    "bits_f32",
    "dt_to_ticks",
    "f32_bits",
    "output_window_stops",
    "seconds_to_ticks",
    "ticks_to_seconds",
    "uint_as_float_bits",
]
