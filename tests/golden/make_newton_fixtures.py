"""Generate the Newton fixed-point fixtures from the reference Namics program.

The reference's own Newton solver (`newton : method : pseudohessian` and
`hessian`, src/sfnewton.cpp:636-747) is run to a tight tolerance from zero
fields on small systems of tests/golden/cases.py; each record holds the
converged free energy, grand potential, per-molecule values, segment
profiles and the reference's iteration count.  Run here (where the
reference binary oracle/_ref/namics is built); writes
tests/golden/newton_fixed_points.json.

Usage: python tests/golden/make_newton_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.cases import CASES  # noqa: E402
from tests.golden.make_fixtures import REF, converged_fixture  # noqa: E402

# (case, method, tolerance)
RUNS = [
    ("cfg1_planar_adsorption", "pseudohessian", 1e-12),
    ("cyl1d_micelle", "pseudohessian", 1e-12),
    ("sph1d_particle", "pseudohessian", 1e-12),
    ("hex_periodic_restricted", "pseudohessian", 1e-12),
    ("planar_markov2", "pseudohessian", 1e-12),
    ("fjc5_planar_adsorption", "pseudohessian", 1e-12),
    ("cyl1d_markov2", "pseudohessian", 1e-12),
    ("cfg1_planar_adsorption", "hessian", 1e-12),
    ("planar2d_particle", "pseudohessian", 1e-11),
]


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    out = []
    for case, method, tol in RUNS:
        try:
            rec = converged_fixture(CASES[case], method, tol, itlim=2000)
        except RuntimeError as e:
            print("skip", case, method, str(e)[-300:], flush=True)
            continue
        rec["case"] = case
        print(case, method, tol, "iterations", rec["iterations"], "residual", rec["residual"], flush=True)
        out.append(rec)
    with open(os.path.join(HERE, "newton_fixed_points.json"), "w") as f:
        json.dump({"runs": out}, f)


if __name__ == "__main__":
    main()
