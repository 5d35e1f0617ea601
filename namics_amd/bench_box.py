"""3-gradient surfactant self-assembly benchmark (the 96^3 box configuration):
a periodic simple-cubic box of 96^3 sites, (A)10(B)10 diblock at total volume
fraction 0.05 in monomeric solvent, chi(B,W) = chi(A,B) = 1.6, stored-subset
propagator memory saving on, starting fields seeded uniform in [-0.1, 0.1]
plus one Gaussian well of width 4 at the box centre; DIIS to 1e-7, FP64, on
the grid path (one launch per phase and per contour step, the iteration
captured in a hipGraph, or launched kernel by kernel with --no-graph).  One
step = one solve from the seeded fields, which stay resident on the device.

    python bench.py --workload box [--steps K] [--warmup W] [--no-graph] [--no-mixed] [--switch S]

The bench line's `mixed_precision` object times the same solves with FP32
propagators while ||g|| is above the switch level (default 1e-4).
"""
import time

import numpy as np


def _solve_timed(s, x0):
    from namics_amd import solve_resident_info
    s.upload(x0)
    t = time.perf_counter()
    info = solve_resident_info([s])[0]
    return info, time.perf_counter() - t


def run_mixed(args, x0, n, f64):
    """The same solves with FP32 propagators while ||g|| > args.switch, then
    FP64 to the tolerance: solves/s, ms per iteration, the FP32 iterations
    and |dF|/F against the FP64 solve."""
    from namics_amd import ScfSystem, synchronize
    from namics_amd.workloads import diblock_box
    s = ScfSystem(diblock_box(n))
    s.set_graph(not args.no_graph)
    s.set_precision(True, args.switch)
    for _ in range(args.warmup):
        _solve_timed(s, x0)
    synchronize()
    t0 = time.perf_counter()
    runs = [_solve_timed(s, x0) for _ in range(args.steps)]
    synchronize()
    elapsed = time.perf_counter() - t0
    info = runs[-1][0]
    res = s.results()
    s.close()
    solve_s = float(np.mean([r[1] for r in runs]))
    return {"value": sum(1 for r in runs if r[0]["status"] == 0) / elapsed, "unit": "solves/s",
            "switch": args.switch, "iterations": info["iterations"], "fp32_iterations": info["fp32_iterations"],
            "ms_per_iteration": 1e3 * solve_s / (info["iterations"] + 1),
            "rel_dF_vs_fp64": abs(res["free_energy"] - f64) / abs(f64)}


def run_box(args, n=96, seed=41):
    from bench import HBM_PEAK_GBS, grid_bytes_per_evaluation
    from namics_amd import ScfSystem, synchronize
    from namics_amd.workloads import diblock_box, seeded_box_fields
    s = ScfSystem(diblock_box(n))
    s.set_graph(not args.no_graph)
    x0 = seeded_box_fields(s.desc["fields"], n, seed)

    def solve():
        return _solve_timed(s, x0)

    for _ in range(args.warmup):
        solve()
    synchronize()
    t0 = time.perf_counter()
    runs = [solve() for _ in range(args.steps)]
    synchronize()
    elapsed = time.perf_counter() - t0
    infos = [r[0] for r in runs]
    conv = sum(1 for i in infos if i["status"] == 0)
    it = infos[-1]["iterations"]
    solve_s = float(np.mean([r[1] for r in runs]))
    ck = s.checkpoint_interval()
    nbytes = sum(grid_bytes_per_evaluation(s.desc, ck, kd=min(k, s.desc["solver"]["m"])) for k in range(it + 1))
    achieved = nbytes / solve_s / 1e9
    res = s.results()
    s.close()
    mixed = None if args.no_mixed else run_mixed(args, x0, n, res["free_energy"])
    return {
        "metric": "converged SCF solves/sec + propagator HBM GB/s at 1e-7 residual; p50/p99 latency",
        "value": conv / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform fields plus a Gaussian well)",
        "config": {"workload": "3-gradient periodic box %d^3, (A)10(B)10 at 0.05, chi_BW = chi_AB = 1.6, "
                               "save_memory (end-point checkpoint interval %d), DIIS 1e-7, FP64, grid path"
                               % (n, ck), "seed": seed,
                   "graph": not args.no_graph},
        "iterations": it,
        "ms_per_iteration": 1e3 * solve_s / (it + 1),
        "latency_ms": {"p50": 1e3 * float(np.percentile([r[1] for r in runs], 50)),
                       "p99": 1e3 * float(np.percentile([r[1] for r in runs], 99))},
        "free_energy": res["free_energy"],
        "mixed_precision": mixed,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole solve (per-kernel durations in profiles/k_step3_box96.md)",
                     "bytes_per_solve": nbytes,
                     "bytes": "every grid-path launch's streams: 9 x 8 B per site and contour step, 3 x 8 B per "
                              "recomputed row, per-site arrays once per evaluation, DIIS history per iteration"},
    }
