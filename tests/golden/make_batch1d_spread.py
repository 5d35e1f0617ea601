"""Rounding-level spread of the DIIS iteration counts of the batched 1-D sample.

The reference's DIIS (src/sfnewton.cpp:1037-1130) can take visibly different
numbers of iterations when its starting fields change in the last bits: for
draw 10 of the sample, starting from 1e-15 instead of 0 takes 73 iterations
instead of 87.  A solve whose arithmetic differs from the reference's only in
rounding (another summation order of a dot product) is therefore held to the
range of iteration counts the reference's own algorithm shows under such
perturbations, not to one trajectory.

The oracle (oracle/liboracle.so) restates the reference bit for bit, including
its DIIS and SVD, and reproduces every fixture's iteration count from zero
fields.  For every draw of batch1d_sample.json that the reference converges,
this script solves from uniform starting fields eps in {0, +-1e-15, +-1e-14,
+-1e-13} and writes the minimum and maximum iteration count into the fixture
(`iterations_spread`).

Usage: python tests/golden/make_batch1d_spread.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.oracle_lib import Oracle  # noqa: E402
from namics_amd.workloads import homopolymer_adsorption  # noqa: E402

EPS = (0.0, 1e-15, -1e-15, 1e-14, -1e-14, 1e-13, -1e-13)


def main():
    path = os.path.join(HERE, "batch1d_sample.json")
    with open(path) as f:
        fx = json.load(f)
    for d in fx["draws"]:
        if "error" in d or d["residual"] > fx["tolerance"]:
            continue
        p = d["params"]
        text = homopolymer_adsorption(p["layers"], p["N"], p["chi"], p["chi_s"], p["phib"], markov=p["markov"],
                                      tol=fx["tolerance"]) + "newton : isaac : iterationlimit : %d\n" % fx[
                                          "iterationlimit"]
        o = Oracle(text)
        its = []
        for eps in EPS:
            x, it, r, st = o.solve(np.full(o.nvar, eps))
            its.append(it if st == 0 else fx["iterationlimit"])
        assert its[0] == d["iterations"], (d["draw"], its[0], d["iterations"])
        d["iterations_spread"] = [min(its), max(its)]
        print("draw", d["draw"], "iterations", its, flush=True)
    fx["spread_eps"] = list(EPS)
    with open(path, "w") as f:
        json.dump(fx, f)


if __name__ == "__main__":
    main()
