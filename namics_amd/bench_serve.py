"""Serving benchmark (the request-stream configuration): Poisson arrivals at a
fixed rate of mixed requests -- 70% 1-gradient adsorption systems drawn as
in the batched sweep, 25% 2-gradient cylindrical brushes (32-128 radial x
32-256 axial layers, N in [50, 400]), 5% 3-gradient diblock boxes
(48^3-96^3) -- each with seeded random parameters and starting fields,
served by the library's request server (state on the device, admission by
HBM footprint, continuous batching).  Reports converged solves/sec over the
whole run and p50/p99 request latency (submission -> end of the request's
own solve), overall and per request class.

    python bench.py --workload serve [--rate R] [--duration S] [--pool-gb G]
"""
import time

import numpy as np

from namics_amd import Server, device_info
from namics_amd.workloads import (batch_draws, brush_cylinder, diblock_box, homopolymer_adsorption,
                                  seeded_box_fields, uniform_fields)


def make_requests(n, seed):
    """n requests: (class, input text, starting fields or None)."""
    rng = np.random.default_rng(seed)
    draws = batch_draws(n, seed=seed + 1)
    out = []
    for i in range(n):
        u = rng.uniform()
        if u < 0.70:
            d = draws[i]
            text = homopolymer_adsorption(d["layers"], d["N"], d["chi"], d["chi_s"], d["phib"], markov=d["markov"],
                                          tol=1e-7)
            text += "newton : isaac : iterationlimit : 1000\n"
            out.append(("1d", text, None))
        elif u < 0.95:
            n_r = int(rng.integers(32, 129))
            n_z = int(rng.integers(32, 257))
            N = int(rng.integers(50, 401))
            chi = float(rng.uniform(0.0, 0.6))
            text = brush_cylinder(chi, n_r=n_r, n_z=n_z, N=N).replace("iterationlimit : 5000",
                                                                           "iterationlimit : 1000")
            out.append(("2d", text, ("uniform", int(rng.integers(1 << 30)), 0.05)))
        else:
            nb = int(rng.integers(48, 97))
            text = diblock_box(nb, itlim=1000)
            out.append(("3d", text, ("box", int(rng.integers(1 << 30)), nb)))
    return out


def fields_for(spec, sysinfo):
    kind = spec[0]
    if kind == "uniform":
        return uniform_fields(sysinfo["nvar"], spec[1], spec[2])
    return seeded_box_fields(sysinfo["fields"], spec[2], spec[1])


def run_serve(args):
    from namics_amd.scf import ScfSystem  # noqa: F401 (library load check)
    cus, hbm = device_info()
    rate = args.rate
    duration = args.duration
    nreq = max(1, int(rate * duration))
    rng = np.random.default_rng(args.seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs = make_requests(nreq, args.seed)
    # starting fields need each request's layout: parse through a throwaway
    # oracle-free description (the server parses again on submission)
    from namics_amd.scf import describe_input
    x0s = []
    for cls, text, spec in reqs:
        if spec is None:
            x0s.append(None)
        else:
            x0s.append(fields_for(spec, describe_input(text)))
    pool = int(args.pool_gb * (1 << 30))
    srv = Server(pool, streams=4)
    ids = {}
    done = {}
    t0 = time.perf_counter()
    k = 0
    while len(done) < nreq:
        now = time.perf_counter() - t0
        while k < nreq and arrivals[k] <= now:
            cls, text, _ = reqs[k]
            ids[srv.submit(text, x0s[k])] = k
            k += 1
        srv.progress()
        for r in srv.results():
            done[ids[r["id"]]] = r
        if k < nreq and len(done) + 64 < k:
            time.sleep(0.0005)
        if now > duration * 20 + 600:
            break
    elapsed = time.perf_counter() - t0
    srv.close()
    res = [done[i] for i in sorted(done)]
    conv = [r for r in res if r["status"] == 0]
    lat = np.array([r["latency_ms"] for r in res])
    by = {}
    for (cls, _, _), r in zip(reqs, res):
        by.setdefault(cls, []).append(r)
    classes = {c: {"requests": len(v), "converged": sum(1 for r in v if r["status"] == 0),
                   "latency_ms_p50": float(np.percentile([r["latency_ms"] for r in v], 50)),
                   "latency_ms_p99": float(np.percentile([r["latency_ms"] for r in v], 99)),
                   "solve_ms_p50": float(np.percentile([r["solve_ms"] for r in v], 50)),
                   "mean_device_mb": float(np.mean([r["device_bytes"] for r in v]) / 2 ** 20)}
               for c, v in sorted(by.items())}
    return {
        "metric": "converged SCF solves/sec + propagator HBM GB/s at 1e-7 residual; p50/p99 latency",
        "value": len(conv) / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1e3 * elapsed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Poisson arrivals, seeded random parameters and fields)",
        "config": {"workload": "request stream: 70% 1-gradient adsorption, 25% 2-gradient brushes "
                               "32-128 x 32-256 N 50-400, 5% 3-gradient diblock boxes 48^3-96^3; "
                               "DIIS 1e-7, iteration limit 1000, FP64",
                   "rate_per_s": rate, "requests": nreq, "pool_gb": args.pool_gb, "seed": args.seed,
                   "compute_units": cus, "hbm_gb": hbm / 2 ** 30},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "queue_ms_p99": float(np.percentile([r["queue_ms"] for r in res], 99)),
        "converged_fraction": len(conv) / max(1, len(res)),
        "classes": classes,
    }
