"""The CPU oracle against the reference's own outputs (golden fixtures).

Fixtures were produced by the reference program compiled from its sources
(tests/golden/make_fixtures.py).  At given fields the oracle restates the
reference's arithmetic in the same order, so agreement is required to a few
ulps; DIIS trajectories must take the same number of iterations.
"""
import numpy as np
import pytest

from tests.helpers import rel_err
from tests.oracle_lib import Oracle

from tests.golden.cases import CASES as _ALL

CASES = list(_ALL)

EVAL_TOL = 1e-12


@pytest.mark.parametrize("case", CASES)
def test_evaluation_matches_reference(golden, case):
    fx = golden[case]
    o = Oracle(fx["text"])
    for ev in fx["evaluation"]:
        g, res = o.residual(np.array(ev["x"]))
        r = o.results()
        assert rel_err(res, ev["residual"]) < EVAL_TOL
        assert rel_err(r["free_energy"], ev["free_energy"]) < EVAL_TOL
        assert rel_err(r["grand_potential"], ev["grand_potential"]) < EVAL_TOL
        for m, d in ev["mol"].items():
            for k in ("GN", "n", "theta", "phibulk", "Mu"):
                assert rel_err(r["mol"][m][k], d[k]) < EVAL_TOL, (m, k)
        for s, p in ev["seg_phi"].items():
            assert rel_err(o.interior(r["seg_phi"][s]), p, 1e-200) < EVAL_TOL, s
        for m, p in ev["mol_phi"].items():
            assert rel_err(o.interior(r["mol_phi"][m]), p, 1e-200) < EVAL_TOL, m


@pytest.mark.parametrize("case", [c for c in CASES])
def test_diis_trajectory_matches_reference(golden, case):
    fx = golden[case]
    for cv in fx["converged"]:
        if cv["method"] != "DIIS":
            continue
        o0 = Oracle(fx["text"])
        o = Oracle(fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : %r\n"
                   "newton : %s : iterationlimit : 5000\n" % (o0.newton_name, o0.newton_name, cv["tolerance"],
                                                              o0.newton_name))
        x0 = np.array(cv["x0"]) if "x0" in cv else np.zeros(o.nvar)
        x, it, res, st = o.solve(x0)
        assert st == 0
        assert it == int(cv["iterations"])
        r = o.results()
        assert rel_err(r["free_energy"], cv["free_energy"]) < 1e-10
        assert rel_err(r["grand_potential"], cv["grand_potential"]) < 1e-10
        for s, p in cv["seg_phi"].items():
            assert rel_err(o.interior(r["seg_phi"][s]), p, 1e-200) < 1e-10


def test_fixed_point_matches_pseudohessian(golden):
    """The reference's Newton (pseudohessian, tolerance 1e-12) and the DIIS
    restatement converged to 1e-11 reach the same fixed point."""
    fx = golden["cfg1_planar_adsorption"]
    cv = [c for c in fx["converged"] if c["method"] == "pseudohessian"][0]
    o0 = Oracle(fx["text"])
    o = Oracle(fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : 1e-11\n"
               "newton : %s : iterationlimit : 5000\n" % (o0.newton_name, o0.newton_name, o0.newton_name))
    x, it, res, st = o.solve(np.zeros(o.nvar))
    assert st == 0
    r = o.results()
    assert rel_err(r["free_energy"], cv["free_energy"]) < 1e-9
    # Omega is a small difference of O(1) terms: compare on the scale of F
    assert abs(r["grand_potential"] - cv["grand_potential"]) < 1e-9 * abs(cv["free_energy"])
    for s, p in cv["seg_phi"].items():
        assert rel_err(o.interior(r["seg_phi"][s]), p, 1e-200) < 1e-8


def _newton_runs():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "newton_fixed_points.json")) as f:
        return json.load(f)["runs"]


@pytest.mark.parametrize("rec", _newton_runs(), ids=lambda r: "%s-%s" % (r["case"], r["method"]))
def test_reference_newton_fixed_points_are_diis_fixed_points(golden, rec):
    """Every fixed point the reference's Newton solver (pseudohessian /
    hessian) reports is the fixed point the DIIS restatement converges to
    from zero fields at 1e-11: free energy and grand potential on the scale
    of F, segment profiles on their own scale, each to what the two
    residuals fix."""
    text = golden[rec["case"]]["text"]
    o0 = Oracle(text)
    o = Oracle(text + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : 1e-11\n"
               "newton : %s : iterationlimit : 20000\n" % (o0.newton_name, o0.newton_name, o0.newton_name))
    x, it, res, st = o.solve(np.zeros(o.nvar))
    assert st == 0
    r = o.results()
    tol = max(1e-9, 100 * rec["residual"])
    scale = abs(rec["free_energy"]) + abs(rec["grand_potential"])
    assert abs(r["free_energy"] - rec["free_energy"]) < tol * scale
    assert abs(r["grand_potential"] - rec["grand_potential"]) < tol * scale
    for s, p in rec["seg_phi"].items():
        a = o.interior(r["seg_phi"][s])
        assert np.max(np.abs(a - np.asarray(p))) < 10 * tol * max(1.0, np.max(np.abs(p))), s
