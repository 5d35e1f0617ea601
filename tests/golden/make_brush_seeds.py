"""Reference runs of the benchmark's brush jobs, one per seed.

The benchmark's 2-gradient workload is a job of three chained DIIS solves of
the cylindrical brush (64 x 128, (X)1(A)199, chi 0 / 0.5 / 0.8, each solve
starting from the previous result, the first from seeded uniform fields in
[-0.05, 0.05]).  This script runs the reference program (oracle/_ref/namics,
built from the reference's sources by oracle/ref.mk) on that job for seeds
FIRST .. FIRST+COUNT-1, the seeds of one benchmark step, and records per
solve: iterations, final residual, free energy and grand potential.

The benchmark and the GPU tests read brush_seeds.json to count the solves
that converge (||g|| <= 1e-7) in both the reference and the library and to
state the library's free-energy error against the reference on them.

Usage: python tests/golden/make_brush_seeds.py [COUNT] [PROCS]
"""
import json
import os
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.make_fixtures import REF, oracle_layout  # noqa: E402
from namics_amd.workloads import (brush_cylinder, brush_job_input, read_job_kal, uniform_fields,  # noqa: E402
                                  write_guess)

FIRST = 3
AMP = 0.05
CHIS = (0.0, 0.5, 0.8)


def run_seed(seed, names, dims, nvar):
    x = uniform_fields(nvar, seed, AMP)
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "output"))
        write_guess(os.path.join(d, "x.ig"), names, dims, x)
        with open(os.path.join(d, "job.in"), "w") as f:
            f.write(brush_job_input("x.ig", CHIS))
        t0 = time.perf_counter()
        r = subprocess.run([REF, "job.in"], cwd=d, capture_output=True, text=True, timeout=3600)
        dt = time.perf_counter() - t0
        kal = os.path.join(d, "output", "job.kal")
        rows = read_job_kal(kal) if os.path.exists(kal) else []
    return {"seed": seed, "rc": r.returncode, "seconds": dt, "solves": rows}


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    count = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    procs = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    o = oracle_layout(brush_cylinder(CHIS[0]))
    names, dims, nvar = o.field_names, o.dims, o.nvar
    out = []
    with ThreadPoolExecutor(procs) as ex:
        for res in ex.map(lambda s: run_seed(s, names, dims, nvar), range(FIRST, FIRST + count)):
            out.append(res)
            print("seed", res["seed"], "rc", res["rc"], "%.1f s" % res["seconds"],
                  [(r["iterations"], "%.1e" % r["residual"]) for r in res["solves"]], flush=True)
    fx = {"first_seed": FIRST, "count": count, "amp": AMP, "chis": list(CHIS), "tolerance": 1e-7,
          "iterationlimit": 5000, "reference": "oracle/_ref/namics (g++ -O2 from the reference sources)",
          "jobs": out}
    with open(os.path.join(HERE, "brush_seeds.json"), "w") as f:
        json.dump(fx, f, indent=0)
    print("wrote brush_seeds.json", flush=True)


if __name__ == "__main__":
    main()
