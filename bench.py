"""Benchmark of the lattice-SCF hot path on one MI355X.

Default workload (the 2-gradient configuration): a polymer brush on a rod in
cylindrical coordinates, 64 radial x 128 axial layers, chains (X)1(A)199
grafted at 0.05 chains per unit rod area, monomeric solvent, solved for
chi(A,W) = 0.0, 0.5, 0.8 in sequence (each solve starts from the previous
one's fields, as the reference chains `start` blocks; the first starts from
seeded uniform random fields in [-0.05, 0.05]), DIIS to ||g|| <= 1e-7
(iteration limit 5000, the input's), FP64.

One job = those three solves for one seed.  One step = one batch of J jobs
(seeds 3 .. 3+J-1, J = 128 by default, 3J solves).  The K timed steps (and,
before them, the W warm-up steps) are handed to the persistent job queue as
one launch of K*J jobs: one workgroup per resident slot (one slot per CU),
each taking the next job when it finishes one, so the stall tail of the
slowest solves is paid once per launch, not once per step.  Fields, seeds
and solver state are resident in HBM before the timed region.

`value` counts the solves that reach ||g|| <= 1e-7 in both this library and
the reference program (tests/golden/brush_seeds.json holds the reference's
own runs of the same 128 seeded jobs), per second of the timed launch.  The
line also states the library's free-energy error against the reference on
those solves, per-solve p50/p99 device latency, the launch's HBM roofline,
the same batch with FP32 propagators and FP64 refinement, one job alone, and
the reference program timed on this host's cores on distinct seeds of the
same batch.

    python bench.py [--steps K] [--warmup W] [--jobs J] [--slots S]
                    [--workload brush|batch1d|serve|box] [--no-cpu] [--no-mixed]
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
TOL = 1e-7
REF_SEEDS = os.path.join(REPO, "tests", "golden", "brush_seeds.json")


def stored_rows(N, ckpt):
    """End-point rows of a chain of length N kept in HBM: all N, or rows
    0, K, 2K, ... with checkpoint interval K > 1."""
    return N if ckpt <= 1 else -(-N // ckpt)


def hbm_bytes_per_evaluation(desc, ckpt=1, w=8, kd=10):
    """Bytes one residual evaluation plus one DIIS update must move through
    HBM on the fused path, each array read or written once (what a perfect
    cache would move).  The propagator recursion itself runs in LDS and
    registers; what must reach HBM is

    * end-point rows: every stored row written once in the forward pass and
      read once in the backward pass, 2 w per stored site and row
      (w = 4 B with FP32 propagators), over the sites the chain kernel stores
      (interior sites on 2-gradient lattices, all sites on 1 gradient);
    * per-site arrays of the evaluation (8 B per site each): fields x read by
      the Boltzmann-weight pass and the residual (2 nit rows), G1 written
      (nseg) and read by every block of the forward and backward passes
      (2 nblocks) and by the density pass (nmm), molecule-segment densities
      written by the chains and read by the density pass (2 nmm), segment
      densities and site averages written and read (4 nseg), the total
      density (2 rows), the residual g written (nit);
    * the DIIS update with kd history entries: xR_k, e_k and x written, x
      and g read, kd entries of both histories read for the Gram row and kd
      of xR for the combination: (3 kd + 5) nvar values;
    * a charged system: the potential block read (fields pass and sweep),
      psi, EE, q, eps written and q, eps read by the sweep and the
      thermodynamics, the sweep's two coefficient rows written and read, its
      g row written: 12 rows."""
    g = desc["gradients"]
    M = desc["M"]
    W = M if g == 1 else desc["MX"] * (desc["MY"] if g >= 2 else 1) * (desc["MZ"] if g == 3 else 1)
    gf = sum(2 * w * m.get("size", 1) * W * stored_rows(m["chainlength"], ckpt) for m in desc["molecules"])
    nit = len(desc["fields"])
    nseg = len(desc["segments"])
    nblocks = sum(m["blocks"] for m in desc["molecules"])
    nmm = sum(m["molmon"] for m in desc["molecules"])
    rows = 3 * nit + nseg + 2 * nblocks + 3 * nmm + 4 * nseg + 2
    if desc.get("charged"):
        rows += 12
    diis = (3 * kd + 5) * desc["nvar"]
    return gf + 8 * M * rows + 8 * diis


def recomputed_rows(N, ckpt):
    """Rows the grid path's backward pass recomputes per evaluation with
    checkpoint interval K: every row between two checkpoints except those of
    the last segment, which the forward pass left in the buffer."""
    if ckpt <= 1:
        return 0
    last = (N - 1) // ckpt * ckpt
    return (N - -(-N // ckpt)) - (N - 1 - last)


def grid_bytes_per_evaluation(desc, ckpt=1, w=8, kd=10):
    """Bytes one residual evaluation plus one DIIS update move on the grid
    path, where every propagation step is a launch that streams its rows
    through memory: forward 3w per site (read G(s-1) and G1, write G(s)),
    backward with composition 6w (read Gb, G1 and Gf, write Gb, read and
    write rho), 3w per recomputed row with checkpoints, over the interior
    sites; the per-site arrays of the other phases and the DIIS history as
    in hbm_bytes_per_evaluation."""
    W = desc["MX"] * max(1, desc["MY"]) * max(1, desc["MZ"])
    chains = sum((9 * m["chainlength"] + 3 * recomputed_rows(m["chainlength"], ckpt)) * w * W
                 for m in desc["molecules"])
    rest = hbm_bytes_per_evaluation(desc, ckpt, w, kd) - \
        sum(2 * w * m.get("size", 1) * W * stored_rows(m["chainlength"], ckpt) for m in desc["molecules"])
    return chains + rest


def host_cores():
    """Cores this process may run on (at most 16, the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def load_reference_seeds():
    """The reference program's own runs of the benchmark jobs: per seed, the
    three solves' iterations, residual and free energy."""
    try:
        with open(REF_SEEDS) as f:
            fx = json.load(f)
    except OSError:
        return {}
    return {j["seed"]: j["solves"] for j in fx["jobs"] if j["rc"] == 0 and len(j["solves"]) == len(CHIS)}


def profiled_traffic(jobs, slots, ckpt):
    """HBM bytes per step of the batched-brush job kernel measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE (copy-calibrated), from the committed
    profile summary, when this run has the profiled configuration."""
    path = os.path.join(REPO, "profiles", "k_fused_jobs_brush128x20.md")
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None, None
    m = re.search(r"traffic_bytes_per_step = ([0-9.eE+-]+)", text)
    c = re.search(r"traffic_config = jobs (\d+), slots (\d+), checkpoint (\d+)", text)
    if m and c and int(c.group(1)) == jobs and int(c.group(2)) == slots and int(c.group(3)) == ckpt:
        return float(m.group(1)), os.path.relpath(path, REPO)
    return None, None


def slot_occupancy(jobs, launch_ms, nslots):
    """How full the persistent launch keeps its slots: busy share of
    slots x launch time, and the time at the end of the launch during which
    fewer than a tenth of the slots still run a job."""
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
            "tail_ms_below_10pct_slots": end - drop}


def compare_with_reference(results, seeds, ref):
    """Solves of `results` (per-job lists of dicts, job j on seed seeds[j])
    whose seed the reference ran: counts converged in the library, in the
    reference and in both, and the relative free-energy difference on the
    solves converged in both."""
    n = conv_lib = conv_ref = both = 0
    rel = []
    for job, seed in zip(results, seeds):
        rs = ref.get(seed)
        if rs is None:
            continue
        for r, q in zip(job, rs):
            n += 1
            a = r["status"] == 0
            b = q["residual"] <= TOL
            conv_lib += a
            conv_ref += b
            if a and b:
                both += 1
                rel.append(abs(r["free_energy"] - q["free_energy"]) / abs(q["free_energy"]))
    return {"solves": n, "converged_library": conv_lib, "converged_reference": conv_ref, "converged_both": both,
            "rel_free_energy_vs_reference": {"max": float(np.max(rel)), "median": float(np.median(rel))}
            if rel else None}


def cpu_baseline_brush(seeds, names, dims, nvar):
    """Time the reference program (oracle/_ref/namics, built from the
    reference's own sources) on the benchmark's jobs: one process per host
    core, each on its own seed of the batch (three start blocks, the first
    from the seeded fields, the rest from the previous result), all started
    together; the throughput counts the converged solves over the wall time
    of the slowest process."""
    ref = os.path.join(REPO, "oracle", "_ref", "namics")
    if not os.path.exists(ref):
        return None
    from namics_amd.workloads import brush_job_input, read_job_kal, uniform_fields, write_guess
    d = tempfile.mkdtemp(prefix="scfbench")
    dirs = []
    for seed in seeds:
        dk = os.path.join(d, "s%d" % seed)
        os.makedirs(os.path.join(dk, "output"))
        write_guess(os.path.join(dk, "x.ig"), names, dims, uniform_fields(nvar, seed, 0.05))
        with open(os.path.join(dk, "job.in"), "w") as f:
            f.write(brush_job_input("x.ig", CHIS))
        dirs.append(dk)
    t0 = time.perf_counter()
    ps = [subprocess.Popen([ref, "job.in"], cwd=dk, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
          for dk in dirs]
    rcs = [p.wait(timeout=900) for p in ps]
    dt = time.perf_counter() - t0
    runs = {}
    for seed, dk in zip(seeds, dirs):
        kal = os.path.join(dk, "output", "job.kal")
        runs[seed] = read_job_kal(kal) if os.path.exists(kal) else []
    return dict(seconds=dt, rc=max(rcs), runs=runs)


def run_brush(args):
    from namics_amd import DeviceBuffer, ScfSystem, autotune_batch, device_info, solve_jobs, synchronize
    from namics_amd.workloads import brush_cylinder, uniform_fields
    ncu = device_info()[0]
    nslots = args.slots if args.slots > 0 else ncu
    J = args.jobs if args.jobs > 0 else 128
    K, W = args.steps, args.warmup
    ref = load_reference_seeds()
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
    ck = args.checkpoint
    if ck == 0 and not args.no_tune:
        tune_batch = autotune_batch([row[0] for row in slots])
        tune = dict(tune or {}, batch=tune_batch)
        ck = tune_batch["chosen"]
    ck = max(1, ck)
    for row in slots:
        for s in row:
            s.set_checkpoint(ck)
    descs = [s.desc for s in slots[0]]
    nvar0 = slots[0][0].nvar
    batch_seeds = list(range(SEED, SEED + J))
    seeds_one = np.stack([uniform_fields(nvar0, s, 0.05) for s in batch_seeds])
    # K steps (and W warm-up steps) of the same batch, one launch each
    seeds_dev = DeviceBuffer(np.tile(seeds_one, (max(K, W, 1), 1)))
    synchronize()
    t_setup = time.perf_counter() - t_setup

    def launch(steps):
        """One persistent launch of `steps` batches; per-job results of the
        last batch, the launch's HIP-event time, and the wall time between
        device synchronisations."""
        synchronize()
        t0 = time.perf_counter()
        res, ms = solve_jobs(slots, seeds_dev, steps * J)
        synchronize()
        return res, ms, time.perf_counter() - t0

    if W > 0:
        launch(W)
    res, launch_ms, elapsed = launch(K)
    solves = [r for job in res for r in job]
    last = res[(K - 1) * J:]
    conv = sum(1 for r in solves if r["status"] == 0)
    cmp = compare_with_reference(last, batch_seeds, ref)
    # every step solves the same batch with the same arithmetic: the count of
    # solves converged in both paths is K times that of one batch
    both = cmp["converged_both"] * K if ref else None
    nbytes = sum(hbm_bytes_per_evaluation(descs[c], ck, kd=min(r["iterations"], descs[c]["solver"]["m"]))
                 * (r["iterations"] + 1) for job in res for c, r in enumerate(job))
    achieved = nbytes / (launch_ms * 1e-3) / 1e9
    out = {
        "metric": "converged SCF solves/sec + propagator HBM GB/s at 1e-7 residual; p50/p99 latency",
        "value": (both if both is not None else conv) / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform random initial fields, one seed per job)",
        "config": {"workload": "configuration 2: 2-gradient cylindrical brush 64x128, (X)1(A)199 grafted "
                               "0.05/area on a rod, chi 0/0.5/0.8 chained (each solve from the previous one's "
                               "fields), DIIS to 1e-7, FP64",
                   "jobs_per_step": J, "solves_per_step": J * len(CHIS), "seeds": [SEED, SEED + J - 1],
                   "slots": nslots, "checkpoint": ck,
                   "launch": "the K steps run as one persistent job-queue launch of K*J jobs"},
        "value_counts": "solves converged (||g|| <= 1e-7) in both the library and the reference program"
                        if both is not None else "solves converged in the library (no reference runs found)",
        "converged_library_per_s": conv / elapsed,
        "vs_reference": cmp if ref else None,
        "latency_ms": {"p50": float(np.percentile([r["solve_ms"] for r in solves], 50)),
                       "p99": float(np.percentile([r["solve_ms"] for r in solves], 99)),
                       "what": "device time of one solve inside the persistent launch"},
        "iterations_p50_p99": [float(np.percentile([r["iterations"] for r in solves], q)) for q in (50, 99)],
        "converged_fraction": conv / max(1, len(solves)),
        "phase_ms_per_iteration": dict(zip(
            ("prepare", "chains", "densities_residual", "diis_other"),
            (np.sum([r["phase_ms"] for r in solves], axis=0) /
             max(1, sum(r["iterations"] + 1 for r in solves))).tolist())),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused_jobs (one launch per timed region)", "avg_launch_ms": launch_ms,
                     "bytes_per_launch": nbytes,
                     "bytes": "HBM bytes the path must move: stored end-point rows written and read once, "
                              "per-site arrays once per evaluation, DIIS history per iteration"},
        "occupancy": slot_occupancy(res, launch_ms, nslots),
        "setup_s": t_setup,
        "autotune": tune,
    }
    traffic, src = profiled_traffic(J, nslots, ck)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic * K
        out["roofline"]["traffic_source"] = src + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated)"
    # one batch with FP32 propagators and FP64 refinement
    if not args.no_mixed:
        if slots[0][0].layout()["strip_sx"] > 0:
            for row in slots:
                for s in row:
                    s.set_precision(True, args.switch)
            resm, msm, elm = launch(1)
            sm = [r for job in resm for r in job]
            convm = sum(1 for r in sm if r["status"] == 0)
            cmpm = compare_with_reference(resm, batch_seeds, ref)
            rel64 = [abs(a["free_energy"] - b["free_energy"]) / abs(b["free_energy"])
                     for ja, jb in zip(resm, last) for a, b in zip(ja, jb) if a["status"] == 0 and b["status"] == 0]
            out["mixed_precision"] = {
                "value": (cmpm["converged_both"] if ref else convm) / elm, "unit": "solves/s",
                "ms_per_step": 1e3 * elm, "steps": 1, "converged_fraction": convm / max(1, len(sm)),
                "switch_residual": args.switch,
                "fp32_iteration_fraction": sum(r["fp32_iterations"] for r in sm) /
                max(1, sum(r["iterations"] for r in sm)),
                "rel_free_energy_vs_fp64": {"max": float(np.max(rel64)), "median": float(np.median(rel64)),
                                            "bar": 1e-8,
                                            "within_bar_fraction": float(np.mean(np.array(rel64) <= 1e-8))}
                if rel64 else None,
                "vs_reference": cmpm if ref else None,
                "latency_ms": {"p50": float(np.percentile([r["solve_ms"] for r in sm], 50)),
                               "p99": float(np.percentile([r["solve_ms"] for r in sm], 99))}}
            for row in slots:
                for s in row:
                    s.set_precision(False)
        else:
            out["mixed_precision"] = {"skipped": "the tuned layout is not a register-strip layout"}
    # one job alone (the latency of the chained job on an otherwise idle GPU),
    # with the end-point storage tuned for one system in flight
    K1 = ck
    if args.checkpoint == 0 and not args.no_tune:
        K1 = autotune_batch(slots[0][:1])["chosen"]
        for s in slots[0]:
            s.set_checkpoint(K1)
    synchronize()
    t1 = time.perf_counter()
    one, _ = solve_jobs(slots[:1], seeds_dev, 1)
    synchronize()
    el1 = time.perf_counter() - t1
    out["single_job"] = {"ms": 1e3 * el1, "solves_per_s": sum(1 for r in one[0] if r["status"] == 0) / el1,
                         "iterations": [r["iterations"] for r in one[0]], "checkpoint": K1,
                         "reference_iterations": [q["iterations"] for q in ref[SEED]] if SEED in ref else None}
    if not args.no_cpu:
        P = host_cores()
        cb = cpu_baseline_brush(batch_seeds[:P], slots[0][0].field_names, slots[0][0].dims, nvar0)
        if cb is not None:
            jobs_cpu = [cb["runs"][s] for s in batch_seeds[:P]]
            conv_cpu = sum(1 for rs in jobs_cpu for q in rs if q["residual"] <= TOL)
            # the same solves in the library (the last timed batch)
            both_cpu = sum(1 for rs, job in zip(jobs_cpu, last[:P]) for q, r in zip(rs, job)
                           if q["residual"] <= TOL and r["status"] == 0)
            out["cpu_baseline"] = {
                "value": both_cpu / cb["seconds"], "unit": "solves/s", "cores": P, "kind": "reference",
                "sample": "the reference program built from its sources, one process per core, each on one job "
                          "of the batch (seeds %d..%d, 3 chained solves each), %.1f s wall; counts solves "
                          "converged in both paths (%d of %d; %d converged in the reference)"
                          % (batch_seeds[0], batch_seeds[P - 1], cb["seconds"], both_cpu, 3 * P, conv_cpu)}
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
    ap.add_argument("--jobs", type=int, default=0, help="brush jobs per step (0: 128, the seeds the reference ran)")
    ap.add_argument("--slots", type=int, default=0, help="resident job slots (0: one per CU)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision batch")
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
