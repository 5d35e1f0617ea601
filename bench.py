"""Benchmark of the lattice-SCF hot path on one MI355X.

Default workload (the 2-gradient configuration): a polymer brush on a rod in
cylindrical coordinates, 64 radial x 128 axial layers, chains (X)1(A)199
grafted at 0.05 chains per unit rod area, monomeric solvent, solved for
chi(A,W) = 0.0, 0.5, 0.8 in sequence (each solve starts from the previous
one's fields, as the reference chains `start` blocks; the first starts from
seeded uniform random fields in [-0.05, 0.05]), DIIS to ||g|| <= 1e-7, FP64.

One job = those three solves.  One step = J independent jobs (one seed
each, J = twice the number of CUs by default) in one launch of the
persistent job queue: one workgroup per resident slot (one slot per CU),
each taking the next job when it finishes one, with fields, seeds and
solver state resident in HBM.  The timed region holds K steps.  Prints one
JSON line: converged solves/sec (FP64), per-solve p50/p99 device latency,
the launch's algorithmic propagator bandwidth against HBM peak, the same
jobs with FP32 propagators and FP64 refinement, one job alone, and the
reference CPU program timed on this host on one job as the baseline.

    python bench.py [--steps K] [--warmup W] [--jobs J] [--slots S] [--workload brush|batch1d|serve]
                    [--no-cpu] [--no-mixed] [--switch R]
"""
import argparse
import json
import os
import re
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


def propagator_bytes(desc, iterations, w=8, ckpt=1):
    """Algorithmic propagator bytes of one solve.  Per residual evaluation a
    first-order (Markov 1) chain of length N streams 3w per site forward (read
    G(s-1), read G1, write G(s)) and 6w backward with composition (read Gb,
    G1, Gf; write Gb; read+write rho): 9 w N per site.  A second-order chain
    with 3 direction components streams 7w forward and 12w backward: 19 w N.
    With save_memory (every ckpt-th end-point row kept), each row not kept
    is recomputed once more: 3w per site for each of N - ceil(N / ckpt)
    rows.  Residual evaluations per solve = iterations + 1."""
    g = desc["gradients"]
    W = desc["MX"] * (desc["MY"] if g >= 2 else 1) * (desc["MZ"] if g == 3 else 1)
    if g == 1:
        W = desc["M"]  # 1-gradient kernels update the ghost sites too
    per_eval = sum((19 if m["markov"] == 2 else 9) * w * m["chainlength"] * W for m in desc["molecules"])
    if ckpt > 1:
        per_eval += sum(3 * w * (m["chainlength"] - -(-m["chainlength"] // ckpt)) * W for m in desc["molecules"])
    return per_eval * (iterations + 1)


def profiled_traffic(jobs, slots, ckpt):
    """HBM bytes per launch of the batched-brush job kernel measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE (copy-calibrated), from the committed
    profile summary, when this run has the profiled configuration."""
    path = os.path.join(REPO, "profiles", "k_fused_jobs_brush512.md")
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None, None
    m = re.search(r"traffic_bytes_per_launch = ([0-9.eE+-]+)", text)
    c = re.search(r"traffic_config = jobs (\d+), slots (\d+), checkpoint (\d+)", text)
    if m and c and int(c.group(1)) == jobs and int(c.group(2)) == slots and int(c.group(3)) == ckpt:
        return float(m.group(1)), os.path.relpath(path, REPO)
    return None, None


def slot_occupancy(jobs, launch_ms, nslots):
    """How full the persistent launch keeps its slots: busy share of
    slots x launch time, the time at the end of the launch during which fewer
    than a tenth of the slots still run a job, and the jobs whose chain has
    more than one solve that stops at the iteration limit."""
    spans = [(job[0]["start_ms"], max(r["start_ms"] + r["solve_ms"] for r in job)) for job in jobs]
    end = max(e for _, e in spans)
    events = sorted([(b, 1) for b, _ in spans] + [(e, -1) for _, e in spans])
    few = max(1, nslots // 10)
    active, drop = 0, 0.0
    for t, d in events:
        if active >= few and active + d < few:
            drop = t
        active += d
    return {"busy_fraction": sum(r["solve_ms"] for job in jobs for r in job) / (nslots * launch_ms),
            "tail_ms_below_10pct_slots": end - drop,
            "jobs_with_2plus_unconverged": sum(1 for job in jobs if sum(r["status"] != 0 for r in job) >= 2)}


def host_cores():
    """Cores this process may run on (at most 16, the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_brush(x0, names, dims, budget_s=60.0, procs=1):
    """Time the reference program (oracle/_ref/namics, built from the
    reference's own sources) on the same job: three start blocks, the first
    from the same seeded fields, the rest from the previous result; `procs`
    copies run at once, one per core, and the result counts all of them over
    the wall time of the slowest."""
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
    dirs = []
    for k in range(procs):
        dk = os.path.join(d, "p%d" % k)
        os.makedirs(os.path.join(dk, "output"))
        with open(os.path.join(dk, "job.in"), "w") as f:
            f.write(text)
        os.link(os.path.join(d, "x.ig"), os.path.join(dk, "x.ig"))
        dirs.append(dk)
    t0 = time.perf_counter()
    ps = [subprocess.Popen([ref, "job.in"], cwd=dk, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
          for dk in dirs]
    rcs = [p.wait(timeout=budget_s * 10) for p in ps]
    dt = time.perf_counter() - t0
    runs = []
    for dk in dirs:
        rows = []
        kal = os.path.join(dk, "output", "job.kal")
        if os.path.exists(kal):
            with open(kal) as f:
                lines = [l.split("\t") for l in f.read().strip().splitlines()]
            rows = [dict(zip(lines[0], l)) for l in lines[1:]]
        runs.append(rows)
    rows = runs[0]
    conv = sum(1 for rs in runs for rw in rs if float(rw["newton_isaac_residual"]) <= 1e-7)
    return dict(seconds=dt, rc=max(rcs), solves=sum(len(rs) for rs in runs), converged=conv, procs=procs,
                iterations=[int(float(rw["newton_isaac_iterations"])) for rw in rows],
                free_energy=[float(rw["sys_NN_free_energy"]) for rw in rows])


def run_brush(args):
    from namics_amd import DeviceBuffer, ScfSystem, autotune_batch, device_info, solve_jobs, synchronize
    from namics_amd.workloads import brush_cylinder, uniform_fields
    ncu = device_info()[0]
    nslots = args.slots if args.slots > 0 else ncu
    J = args.jobs if args.jobs > 0 else 2 * nslots
    t_setup = time.perf_counter()
    # launch layout: tuned once on one system of the brush's shape; every
    # slot system built afterwards starts from the tuned layout
    tune = None
    if not args.no_tune:
        probe = ScfSystem(brush_cylinder(CHIS[-1]))
        probe.upload(uniform_fields(probe.nvar, SEED, 0.05))
        tune = probe.autotune(evals=5)
        tune["layout"] = probe.layout()
        probe.close()
    slots = [[ScfSystem(brush_cylinder(c)) for c in CHIS] for _ in range(nslots)]
    # end-point storage: tuned with all slots in flight (every row, or every
    # other row with the odd rows recomputed on chip), unless given
    K = args.checkpoint
    if K == 0 and not args.no_tune:
        tune_batch = autotune_batch([row[0] for row in slots])
        tune = dict(tune or {}, batch=tune_batch)
        K = tune_batch["chosen"]
    K = max(1, K)
    for row in slots:
        for s in row:
            s.set_checkpoint(K)
    descs = [s.desc for s in slots[0]]
    nvar0 = slots[0][0].nvar
    seeds = np.stack([uniform_fields(nvar0, SEED + j, 0.05) for j in range(J)])
    seeds_dev = DeviceBuffer(seeds)
    synchronize()
    t_setup = time.perf_counter() - t_setup

    def run(steps, warmup):
        for _ in range(warmup):
            solve_jobs(slots, seeds_dev, J)
        synchronize()
        t0 = time.perf_counter()
        rec = [solve_jobs(slots, seeds_dev, J) for _ in range(steps)]
        synchronize()
        return rec, time.perf_counter() - t0

    def summary(rec, elapsed):
        solves = [r for res, _ in rec for job in res for r in job]
        conv = sum(1 for r in solves if r["status"] == 0)
        launch_ms = float(np.mean([ms for _, ms in rec]))
        nbytes = float(np.mean([sum(propagator_bytes(descs[c], job[c]["iterations"], ckpt=K)
                                    for job in res for c in range(len(CHIS))) for res, _ in rec]))
        return solves, conv, launch_ms, nbytes

    rec, elapsed = run(args.steps, args.warmup)
    solves, converged, launch_ms, nbytes = summary(rec, elapsed)
    achieved = nbytes / (launch_ms * 1e-3) / 1e9
    F64 = [[r["free_energy"] for r in job] for job in rec[-1][0]]
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
        "data": "synthetic (seeded uniform random initial fields, one seed per job)",
        "config": {"workload": "2-gradient cylindrical brush 64x128, (X)1(A)199 grafted 0.05/area on a rod, "
                               "chi 0/0.5/0.8 chained (each solve starts from the previous one's fields), "
                               "DIIS 1e-7, FP64",
                   "jobs_per_step": J, "solves_per_step": J * len(CHIS), "seeds": [SEED, SEED + J - 1],
                   "slots": nslots, "checkpoint": K},
        "latency_ms": {"p50": float(np.percentile([r["solve_ms"] for r in solves], 50)),
                       "p99": float(np.percentile([r["solve_ms"] for r in solves], 99)),
                       "what": "device time of one solve inside the persistent launch"},
        "iterations_p50_p99": [float(np.percentile([r["iterations"] for r in solves], q)) for q in (50, 99)],
        "converged_fraction": converged / max(1, len(solves)),
        # device time per DIIS iteration by phase, over every solve of the
        # timed steps (device clock inside the persistent launch)
        "phase_ms_per_iteration": dict(zip(
            ("prepare", "chains", "densities_residual", "diis_other"),
            (np.sum([r["phase_ms"] for r in solves], axis=0) /
             max(1, sum(r["iterations"] + 1 for r in solves))).tolist())),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused_jobs (one launch per step)", "avg_launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": nbytes},
        "occupancy": slot_occupancy(rec[-1][0], rec[-1][1], nslots),
        "setup_s": t_setup,
        "autotune": tune,
    }
    traffic, src = profiled_traffic(J, nslots, K)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_source"] = src + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated)"
        out["roofline"]["hbm_gbs_measured_traffic"] = traffic / (launch_ms * 1e-3) / 1e9
        # measured HBM bytes over the launch against the peak (frac counts the
        # algorithmic bytes, most of which stay on chip)
        out["roofline"]["frac_measured_traffic"] = out["roofline"]["hbm_gbs_measured_traffic"] / HBM_PEAK_GBS
    # the same jobs with FP32 propagators and FP64 refinement
    mixed_ok = not args.no_mixed and slots[0][0].layout()["strip_sx"] > 0
    if not args.no_mixed and not mixed_ok:
        out["mixed_precision"] = {"skipped": "the tuned layout is not a register-strip layout"}
    if mixed_ok:
        for row in slots:
            for s in row:
                s.set_precision(True, args.switch)
        recm, elm = run(args.steps, args.warmup)
        sm, convm, msm, _ = summary(recm, elm)
        Fm = [[r["free_energy"] for r in job] for job in recm[-1][0]]
        both = [(a, b) for ja, jb, jfa, jfb in zip(Fm, F64, recm[-1][0], rec[-1][0])
                for a, b, ra, rb in zip(ja, jb, jfa, jfb) if ra["status"] == 0 and rb["status"] == 0]
        out["mixed_precision"] = {
            "value": convm / elm, "unit": "solves/s", "ms_per_step": 1e3 * elm / args.steps,
            "converged_fraction": convm / max(1, len(sm)), "switch_residual": args.switch,
            "fp32_iteration_fraction": sum(r["fp32_iterations"] for r in sm) / max(1, sum(r["iterations"] for r in sm)),
            "median_rel_free_energy_diff_vs_fp64": float(np.median([abs(a - b) / abs(b) for a, b in both]))
            if both else None,
            "latency_ms": {"p50": float(np.percentile([r["solve_ms"] for r in sm], 50)),
                           "p99": float(np.percentile([r["solve_ms"] for r in sm], 99))}}
        for row in slots:
            for s in row:
                s.set_precision(False)
    # one job alone (the latency of the chained job on an otherwise idle GPU),
    # with the end-point storage tuned for one system in flight
    K1 = K
    if args.checkpoint == 0 and not args.no_tune:
        K1 = autotune_batch(slots[0][:1])["chosen"]
        for s in slots[0]:
            s.set_checkpoint(K1)
    one, el1 = None, 0.0
    t1 = time.perf_counter()
    one, _ = solve_jobs(slots[:1], seeds_dev, 1)
    synchronize()
    el1 = time.perf_counter() - t1
    out["single_job"] = {"ms": 1e3 * el1, "solves_per_s": sum(1 for r in one[0] if r["status"] == 0) / el1,
                         "iterations": [r["iterations"] for r in one[0]], "checkpoint": K1}
    if not args.no_cpu:
        P = host_cores()
        cb = cpu_baseline_brush(seeds[0], slots[0][0].field_names, slots[0][0].dims, procs=P)
        if cb is not None:
            out["cpu_baseline"] = {"value": cb["converged"] / cb["seconds"], "unit": "solves/s", "cores": P,
                                   "kind": "reference",
                                   "sample": "%d copies of one job (3 chained solves, seed %d) by the reference "
                                             "program built from its sources, one process per core, %.1f s wall; "
                                             "iterations %s" % (P, SEED, cb["seconds"], cb["iterations"]),
                                   "free_energy": cb["free_energy"]}
    for row in slots:
        for s in row:
            s.close()
    seeds_dev.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="brush")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--jobs", type=int, default=0, help="independent brush jobs per step (0: twice the slots)")
    ap.add_argument("--slots", type=int, default=0, help="resident job slots (0: one per CU)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision rerun")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch layout (no autotuning)")
    ap.add_argument("--no-graph", action="store_true", help="box: launch the grid-path kernels one by one")
    ap.add_argument("--checkpoint", type=int, default=0,
                    help="brush: keep every K-th end-point row (save_memory); 0 tunes it with all slots in flight")
    ap.add_argument("--switch", type=float, default=1e-4, help="mixed precision: FP64 below this residual")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: requests per second")
    ap.add_argument("--duration", type=float, default=20.0, help="serve: seconds of arrivals")
    ap.add_argument("--pool-gb", type=float, default=64.0, help="serve: HBM pool for requests")
    ap.add_argument("--seed", type=int, default=7, help="serve: request-stream seed")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(independent systems shard across GPUs with no communication; run one process per GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "brush":
        out = run_brush(args)
    elif args.workload == "serve":
        from namics_amd.bench_serve import run_serve
        out = run_serve(args)
    elif args.workload == "box":
        from namics_amd.bench_box import run_box
        out = run_box(args)
    else:
        from namics_amd.bench_batch import run_batch
        out = run_batch(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
