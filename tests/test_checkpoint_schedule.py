"""The adjoint's binomial checkpoint schedule (host side of the HIP library; no GPU needed).

``inu_adj_schedule_cost`` builds the optimal offline schedule for reversing n
segments with s checkpoint slots (slot 0 holding the start) and returns how
many segments it advances without tape. Griewank's binomial result fixes
that number: with beta(c, r) = C(c + r, c) and r the repetition number,
beta(s, r - 1) < n <= beta(s, r), the fewest advances are
r n - beta(s + 1, r - 1) (Griewank and Walther, "Algorithm 799: revolve",
ACM TOMS 26(1), 2000). Two special cases also follow from first principles:
one slot advances 0 + 1 + ... + (n - 1) segments, and with a slot for every
segment each segment is advanced once, except the last.
"""

from __future__ import annotations

from math import comb

import pytest

from inundation_amd.hip import lib as L


def _beta(c: int, r: int) -> int:
    return comb(c + r, c)


def _griewank(n: int, slots: int) -> int:
    if n <= 1:
        return 0
    r = 0
    while _beta(slots, r) < n:
        r += 1
    return r * n - _beta(slots + 1, r - 1)


@pytest.mark.parametrize("slots", [1, 2, 3, 4, 5, 8, 13, 20])
def test_schedule_cost_is_griewanks_binomial_optimum(slots: int) -> None:
    for n in list(range(1, 200)) + [500, 1000, 2048]:
        assert L.schedule_cost(n, slots) == _griewank(n, slots), (n, slots)


def test_special_cases() -> None:
    for n in (1, 2, 7, 40):
        assert L.schedule_cost(n, 1) == n * (n - 1) // 2
        assert L.schedule_cost(n, 64) == n - 1
    assert L.schedule_cost(64, 64) == 63
    assert L.schedule_cost(2048, 64) > 2047


def test_more_slots_never_cost_more() -> None:
    for n in (10, 100, 1000):
        costs = [L.schedule_cost(n, s) for s in range(1, 25)]
        assert all(a >= b for a, b in zip(costs, costs[1:])), (n, costs)


def test_out_of_range_requests_are_refused() -> None:
    for n, s in ((0, 3), (10, 0), (10, 65), (2049, 4)):
        with pytest.raises(ValueError):
            L.schedule_cost(n, s)
