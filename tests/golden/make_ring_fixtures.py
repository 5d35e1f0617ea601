"""Reference fixtures for ring molecules
(tests/golden/ring_cases.py).

Runs the reference program (oracle/_ref/namics, built from its sources by
oracle/ref.mk) on every ring case: the residual evaluation at seeded
fields and the DIIS solve from zero fields, with F, Omega, per-molecule GN,
n, theta, phibulk, mu and every density profile (kal / pro files), as
make_fixtures.py records them for the linear cases.

Usage: python tests/golden/make_ring_fixtures.py [case ...]
       (writes tests/golden/ring/<case>.json)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from tests.golden.ring_cases import RING  # noqa: E402
from tests.golden.make_fixtures import REF, converged_fixture, evaluation_fixture  # noqa: E402


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    only = set(sys.argv[1:])
    out_dir = os.path.join(HERE, "ring")
    os.makedirs(out_dir, exist_ok=True)
    for name, case in RING.items():
        if only and name not in only:
            continue
        fx = {"case": name, "text": case["text"], "evaluation": [], "converged": []}
        for seed, amp in case.get("eval", []):
            fx["evaluation"].append(evaluation_fixture(case, seed, amp))
        for conv in case.get("converged", []):
            fx["converged"].append(converged_fixture(case, **conv))
        with open(os.path.join(out_dir, name + ".json"), "w") as f:
            json.dump(fx, f)
        print("wrote", name, [e["mol"] for e in fx["evaluation"]], flush=True)


if __name__ == "__main__":
    main()
