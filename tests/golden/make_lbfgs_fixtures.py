"""Reference fixtures for the LBFGS solver (`newton : method : LBFGS`).

Runs the reference program (oracle/_ref/namics, built from its sources by
oracle/ref.mk) with method LBFGS from seeded uniform starting fields, on a
few systems of the golden set (1-D first and second order, curved 1-D, 2-D
strip and hexagonal, 3-D periodic and mirror), and records per system: the
LBFGS iteration count and final ||g|| from the reference's own report
("Problem solved: K iterations, |g|: R" -- its kal `residual` is not updated
on this path), and the converged F, Omega, per-molecule GN, n, theta,
phibulk, mu and density profiles (kal / pro files).

Usage: python tests/golden/make_lbfgs_fixtures.py   (writes lbfgs_runs.json)
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

RUNS = [("cfg1_planar_adsorption", 61), ("planar_markov2", 62), ("cyl1d_micelle", 63), ("brush2d_rod_strip", 64),
        ("hex2d_cyl_strip", 65), ("diblock3d_periodic", 66), ("particle3d_mirror", 67)]
AMP, TOL, ITLIM = 0.05, 1e-9, 3000


def lbfgs_run(case, seed):
    o = oracle_layout(case["text"])
    x = np.random.default_rng(seed).uniform(-AMP, AMP, o.nvar)
    with tempfile.TemporaryDirectory() as d:
        write_guess(os.path.join(d, "x.ig"), o.field_names, o.dims, x, fjc=o.fjc)
        text = case["text"] + "\n" + output_lines(o) + (
            "sys : %s : initial_guess : file\n"
            "sys : %s : guess_inputfile : x.ig\n"
            "newton : %s : method : LBFGS\n"
            "newton : %s : tolerance : %r\n"
            "newton : %s : iterationlimit : %d\n"
            "newton : %s : e_info : false\n"
            "start\n") % (o.sys_name, o.sys_name, o.newton_name, o.newton_name, TOL, o.newton_name, ITLIM,
                          o.newton_name)
        r = run_ref(d, "lb", text)
        if r.returncode != 0:
            raise RuntimeError("reference failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        m = re.search(r"Problem solved: (\d+) iterations,\s+\|g\|: (\S+)", r.stdout)
        kal = read_kal(os.path.join(d, "output", "lb.kal"))
        pro = read_pro(os.path.join(d, "output", "lb.pro"), o.gradients)
    out = collect(o, kal, pro)
    out["iterations"] = int(m.group(1))
    out["residual"] = float(m.group(2))
    out["seed"] = seed
    out["amp"] = AMP
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    fx = {"tolerance": TOL, "iterationlimit": ITLIM, "runs": {}}
    for name, seed in RUNS:
        r = lbfgs_run(CASES[name], seed)
        print(name, "iterations", r["iterations"], "|g|", r["residual"], "F", r["free_energy"], flush=True)
        fx["runs"][name] = r
    with open(os.path.join(HERE, "lbfgs_runs.json"), "w") as f:
        json.dump(fx, f)


if __name__ == "__main__":
    main()
