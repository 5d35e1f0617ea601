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


def test_occupancy_without_tail():
    jobs = [[_solve(0, 5)] for _ in range(10)]
    occ = bench.slot_occupancy(jobs, 5.0, 10)
    assert occ["busy_fraction"] == 1.0
    assert occ["tail_ms_below_10pct_slots"] == 0.0


def _desc(N=200, ckpt_sites=(64, 128), nvar=25740):
    return {"gradients": 2, "M": 66 * 130, "MX": ckpt_sites[0], "MY": ckpt_sites[1], "MZ": 0, "nvar": nvar,
            "fields": ["X", "A", "W"], "segments": [{}] * 4,
            "molecules": [{"chainlength": N, "size": 1, "blocks": 2, "molmon": 2},
                          {"chainlength": 1, "size": 1, "blocks": 1, "molmon": 1}],
            "solver": {"m": 10}}


def test_hbm_bytes_count_stored_rows_once_each_way():
    d = _desc()
    full, half = bench.hbm_bytes_per_evaluation(d, 1), bench.hbm_bytes_per_evaluation(d, 2)
    # interval 2 keeps rows 0, 2, ..., 198 of the 200-segment chain and the
    # solvent's single row: 100 rows fewer, each 8 B x 8192 sites written and read
    assert full - half == 100 * 2 * 8 * 8192
    # FP32 end-point rows halve the row bytes only
    f32 = bench.hbm_bytes_per_evaluation(d, 1, w=4)
    assert full - f32 == 201 * 2 * 4 * 8192
    # each DIIS history entry adds three reads of nvar doubles
    assert bench.hbm_bytes_per_evaluation(d, 1, kd=3) - bench.hbm_bytes_per_evaluation(d, 1, kd=2) == 3 * 8 * 25740


def test_compare_with_reference_counts_both_and_free_energy():
    ref = {3: [{"residual": 5e-8, "free_energy": -2.0}, {"residual": 1e-3, "free_energy": 1.0},
               {"residual": 9e-8, "free_energy": 4.0}]}
    res = [[{"status": 0, "free_energy": -2.0 * (1 + 1e-9)}, {"status": 0, "free_energy": 1.0},
            {"status": 1, "free_energy": 4.0}],
           [{"status": 0, "free_energy": 0.0}] * 3]   # seed 4: no reference run, not counted
    c = bench.compare_with_reference(res, [3, 4], ref)
    assert (c["solves"], c["converged_library"], c["converged_reference"], c["converged_both"]) == (3, 2, 2, 1)
    assert abs(c["rel_free_energy_vs_reference"]["max"] - 1e-9) < 1e-15
