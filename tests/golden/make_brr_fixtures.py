"""Reference fixtures for the Broyden rank-reduction solver (`newton : method : BRR`).

Runs the reference program (oracle/_ref/namics, built from its sources by
oracle/ref.mk) with method BRR from seeded uniform starting fields on a few
systems of the golden set, with e_info on and i_info 1, and records per
system: the number of BRR iterations (the reference prints one line per
iteration, "i = it  g = ||g||  alpha = ...", it counting from 0, and
"BRR: that will do. Error: ..." when it stops at the tolerance), the
range of iteration counts the reference takes from the same fields perturbed
by 1e-15 (relative, 16 draws), the final ||g||, and the converged F, Omega, per-molecule GN, n, theta, phibulk, mu
and density profiles (kal / pro files).

Usage: python tests/golden/make_brr_fixtures.py   (writes brr_runs.json)
"""
import json
import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.cases import CASES  # noqa: E402
from tests.golden.make_fixtures import (REF, collect, oracle_layout, output_lines, read_kal, read_pro,  # noqa: E402
                                        run_ref)
from namics_amd.workloads import write_guess  # noqa: E402

# (case, seed, amplitude): runs the reference's BRR converges (it diverges
# from some starting fields, e.g. the rod brush from seeded ones, where its
# partition function underflows and the program stops)
RUNS = [("cfg1_planar_adsorption", 81, 0.05), ("cyl1d_micelle", 82, 0.05), ("planar_markov2", 81, 0.05),
        ("brush2d_rod_strip", 81, 0.0), ("diblock3d_periodic", 81, 0.05)]
TOL, ITLIM = 1e-7, 3000


SPREAD = 16  # relative perturbations of 1e-15 for the iteration spread


def brr_run(case, seed, amp, perturb=None):
    o = oracle_layout(case["text"])
    x = np.random.default_rng(seed).uniform(-amp, amp, o.nvar)
    if perturb is not None:
        x = x * (1 + 1e-15 * np.random.default_rng(perturb).standard_normal(o.nvar)) + \
            (1e-15 * np.random.default_rng(perturb).standard_normal(o.nvar) if amp == 0 else 0)
    with tempfile.TemporaryDirectory() as d:
        write_guess(os.path.join(d, "x.ig"), o.field_names, o.dims, x, fjc=o.fjc)
        text = case["text"] + "\n" + output_lines(o) + (
            "sys : %s : initial_guess : file\n"
            "sys : %s : guess_inputfile : x.ig\n"
            "newton : %s : method : BRR\n"
            "newton : %s : tolerance : %r\n"
            "newton : %s : iterationlimit : %d\n"
            "newton : %s : e_info : true\n"
            "newton : %s : i_info : 1\n"
            "start\n") % (o.sys_name, o.sys_name, o.newton_name, o.newton_name, TOL, o.newton_name, ITLIM,
                          o.newton_name, o.newton_name)
        r = run_ref(d, "br", text)
        if r.returncode != 0:
            raise RuntimeError("reference failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        its = [int(v) for v in re.findall(r"^i = (\d+)\s+g = ", r.stdout, re.M)]
        done = re.search(r"BRR: that will do\. Error: (\S+)", r.stdout)
        kal = read_kal(os.path.join(d, "output", "br.kal"))
        pro = read_pro(os.path.join(d, "output", "br.pro"), o.gradients)
    out = collect(o, kal, pro)
    out["iterations"] = (its[-1] + 1) if its else 0
    out["converged"] = done is not None
    out["error"] = float(done.group(1)) if done else None
    out["seed"] = seed
    out["amp"] = amp
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    fx = {"tolerance": TOL, "iterationlimit": ITLIM, "runs": {}}
    for name, seed, amp in RUNS:
        r = brr_run(CASES[name], seed, amp)
        # the reference's own iteration count from starting fields that
        # differ in the last bits: BRR's trajectory is decided at the
        # rounding level, so a restatement is held to this range
        its = [r["iterations"]]
        for k in range(SPREAD):
            q = brr_run(CASES[name], seed, amp, perturb=1000 + k)
            if q["converged"]:
                its.append(q["iterations"])
        r["iterations_spread"] = [min(its), max(its)]
        print(name, "iterations", r["iterations"], "spread", r["iterations_spread"], "converged", r["converged"],
              "error", r["error"], "F", r["free_energy"], flush=True)
        fx["runs"][name] = r
    with open(os.path.join(HERE, "brr_runs.json"), "w") as f:
        json.dump(fx, f)


if __name__ == "__main__":
    main()
