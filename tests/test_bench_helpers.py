"""Host-side helpers of bench.py (no GPU): the slot-occupancy figures of a
persistent launch, computed from per-solve start and duration."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def _solve(start, ms, status=0):
    return {"start_ms": start, "solve_ms": ms, "status": status}


def test_occupancy_of_a_known_timeline():
    # 20 slots; 19 one-solve-per-step jobs end at t = 3, one job runs to t = 15
    # with two solves stopped at the iteration limit
    long_job = [_solve(0, 1), _solve(1, 9, 1), _solve(10, 5, 1)]
    short = [[_solve(0, 1), _solve(1, 1), _solve(2, 1)] for _ in range(19)]
    occ = bench.slot_occupancy([long_job] + short, 15.0, 20)
    assert occ["busy_fraction"] == (15 + 19 * 3) / (20 * 15.0)
    assert occ["tail_ms_below_10pct_slots"] == 12.0   # below 2 active from t = 3 to 15
    assert occ["jobs_with_2plus_unconverged"] == 1


def test_occupancy_without_tail():
    jobs = [[_solve(0, 5)] for _ in range(10)]
    occ = bench.slot_occupancy(jobs, 5.0, 10)
    assert occ["busy_fraction"] == 1.0
    assert occ["tail_ms_below_10pct_slots"] == 0.0
    assert occ["jobs_with_2plus_unconverged"] == 0
