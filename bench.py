"""Benchmark of the lattice-SCF hot path on one MI355X.

Default workload (the 2-gradient configuration): a polymer brush on a rod in
cylindrical coordinates, 64 radial x 128 axial layers, chains (X)1(A)199
grafted at 0.05 chains per unit rod area, monomeric solvent, solved for
chi(A,W) = 0.0, 0.5, 0.8 in sequence (each solve starts from the previous
one's fields, as the reference chains `start` blocks; the first starts from
seeded uniform random fields in [-0.05, 0.05]), DIIS to ||g|| <= 1e-7, FP64.

One step = the whole job (three solves).  The timed region holds K steps
with fields and state already resident in HBM.  Prints one JSON line with
solves/sec, per-solve p50/p99 latency, the fused kernel's algorithmic
propagator bandwidth against HBM peak, and the reference CPU program timed
on this host as the baseline.

    python bench.py [--steps K] [--warmup W] [--workload brush|batch1d] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
SEED = 3
CHIS = (0.0, 0.5, 0.8)


def propagator_bytes(desc, iterations, w=8):
    """Algorithmic propagator bytes of one solve.  Per residual evaluation a
    first-order (Markov 1) chain of length N streams 3w per site forward (read
    G(s-1), read G1, write G(s)) and 6w backward with composition (read Gb,
    G1, Gf; write Gb; read+write rho): 9 w N per site.  A second-order chain
    with 3 direction components streams 7w forward and 12w backward: 19 w N.
    Residual evaluations per solve = iterations + 1."""
    g = desc["gradients"]
    W = desc["MX"] * (desc["MY"] if g >= 2 else 1) * (desc["MZ"] if g == 3 else 1)
    if g == 1:
        W = desc["M"]  # 1-gradient kernels update the ghost sites too
    per_eval = sum((19 if m["markov"] == 2 else 9) * w * m["chainlength"] * W for m in desc["molecules"])
    return per_eval * (iterations + 1)


def cpu_baseline_brush(x0, names, dims, budget_s=60.0):
    """Time the reference program (oracle/_ref/namics, built from the
    reference's own sources) on the same job: three start blocks, the first
    from the same seeded fields, the rest from the previous result."""
    ref = os.path.join(REPO, "oracle", "_ref", "namics")
    if not os.path.exists(ref):
        return None
    from namics_amd.workloads import brush_cylinder, write_guess
    d = tempfile.mkdtemp(prefix="scfbench")
    os.makedirs(os.path.join(d, "output"))
    write_guess(os.path.join(d, "x.ig"), names, dims, x0)
    text = brush_cylinder(CHIS[0])
    text += ("sys : NN : initial_guess : file\nsys : NN : guess_inputfile : x.ig\n"
             "newton : isaac : e_info : false\n"
             "output : kal : append : true\nkal : newton : isaac : iterations\n"
             "kal : newton : isaac : residual\nkal : sys : NN : free_energy\nstart\n")
    for c in CHIS[1:]:
        text += "mon : A : chi_W : %r\nsys : NN : initial_guess : previous_result\nstart\n" % c
    with open(os.path.join(d, "job.in"), "w") as f:
        f.write(text)
    t0 = time.perf_counter()
    r = subprocess.run([ref, "job.in"], cwd=d, capture_output=True, text=True, timeout=budget_s * 10)
    dt = time.perf_counter() - t0
    rows = []
    kal = os.path.join(d, "output", "job.kal")
    if os.path.exists(kal):
        with open(kal) as f:
            lines = [l.split("\t") for l in f.read().strip().splitlines()]
        rows = [dict(zip(lines[0], l)) for l in lines[1:]]
    conv = sum(1 for rw in rows if float(rw["newton_isaac_residual"]) <= 1e-7)
    return dict(seconds=dt, rc=r.returncode, solves=len(rows), converged=conv,
                iterations=[int(float(rw["newton_isaac_iterations"])) for rw in rows],
                free_energy=[float(rw["sys_NN_free_energy"]) for rw in rows])


def run_brush(args):
    import torch
    from namics_amd import ScfSystem, solve_resident
    from namics_amd.workloads import brush_cylinder, uniform_fields

    systems = [ScfSystem(brush_cylinder(c)) for c in CHIS]
    seed_sys = ScfSystem(brush_cylinder(CHIS[0]))
    x0 = uniform_fields(seed_sys.nvar, SEED, 0.05)
    seed_sys.upload(x0)
    torch.cuda.synchronize()

    def step(record):
        systems[0].carry_from(seed_sys)
        out = []
        for i, s in enumerate(systems):
            if i > 0:
                s.carry_from(systems[i - 1])
            it, res, st, kms = solve_resident([s])[0]
            out.append((it, res, st, kms))
        record.append(out)

    warm = []
    for _ in range(args.warmup):
        step(warm)
    rec = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        step(rec)
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    solves = [s for stp in rec for s in stp]
    converged = sum(1 for s in solves if s[2] == 0)
    kernel_ms = [s[3] for s in solves]
    nbytes = [propagator_bytes(systems[i].desc, rec[k][i][0]) for k in range(len(rec)) for i in range(3)]
    achieved = sum(nbytes) / (sum(kernel_ms) * 1e-3) / 1e9
    results = [s.results() for s in systems]
    out = {
        "metric": "converged SCF solves/sec + propagator HBM GB/s at 1e-7 residual; p50/p99 latency",
        "value": converged / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform random initial fields)",
        "config": {"workload": "2-gradient cylindrical brush 64x128, (X)1(A)199 grafted 0.05/area on a rod, "
                               "chi 0/0.5/0.8 chained, DIIS 1e-7, FP64",
                   "solves_per_step": 3, "seed": SEED},
        "latency_ms": {"p50": float(np.percentile(kernel_ms, 50)), "p99": float(np.percentile(kernel_ms, 99)),
                       "job_p50": 1e3 * float(np.percentile(lat, 50)), "job_p99": 1e3 * float(np.percentile(lat, 99))},
        "iterations": [s[0] for s in rec[-1]],
        "converged_fraction": converged / max(1, len(solves)),
        "free_energy": [r["free_energy"] for r in results],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused", "avg_launch_ms": float(np.mean(kernel_ms)),
                     "algorithmic_bytes_per_launch": float(np.mean(nbytes))},
    }
    if not args.no_cpu:
        cb = cpu_baseline_brush(x0, seed_sys.field_names, seed_sys.dims)
        if cb is not None:
            out["cpu_baseline"] = {"value": cb["converged"] / cb["seconds"], "unit": "solves/s", "cores": 1,
                                   "kind": "reference",
                                   "sample": "the full job (3 chained solves) by the reference program built from "
                                             "its sources, one core, %.1f s; iterations %s" % (cb["seconds"],
                                                                                               cb["iterations"]),
                                   "free_energy": cb["free_energy"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="brush")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent systems shard trivially across GPUs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "brush":
        out = run_brush(args)
    else:
        from namics_amd.bench_batch import run_batch
        out = run_batch(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
