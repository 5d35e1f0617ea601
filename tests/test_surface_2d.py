"""Surface boundaries on 2-gradient lattices against the reference's own
outputs.

A surface ghost keeps its own value (set_bounds copies nothing there), so
the values the reference computes on it over its flattened arrays (its
propagator passes are shifted copies over the index range, which wrap
across rows) reach the interior.  The oracle restates those passes and
matches the reference fixtures (tests/golden/make_surface_fixtures.py) to a
few ulps, with the same DIIS iteration counts; the library runs these
systems with the chains over every site in the same array form (one
workgroup per system) and matches the fixtures to 1e-10 at given fields and
at the converged state.
"""
import json
import os

import numpy as np
import pytest

from tests.golden.surface_cases import SURFACE
from tests.helpers import rel_err, scaled_err
from tests.oracle_lib import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(name):
    with open(os.path.join(HERE, "golden", "surface", name + ".json")) as f:
        return json.load(f)


def _solver_text(fx, o, tol):
    return fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : %r\n" \
        "newton : %s : iterationlimit : 5000\n" % (o.newton_name, o.newton_name, tol, o.newton_name)


@pytest.mark.parametrize("case", list(SURFACE))
def test_oracle_surface_matches_reference(case):
    fx = _fixture(case)
    o = Oracle(fx["text"])
    for ev in fx["evaluation"]:
        g, res = o.residual(np.array(ev["x"]))
        r = o.results()
        assert rel_err(res, ev["residual"]) < 1e-12
        assert rel_err(r["free_energy"], ev["free_energy"]) < 1e-12
        for s, p in ev["seg_phi"].items():
            assert rel_err(o.interior(r["seg_phi"][s]), p, 1e-200) < 1e-12, s
    for cv in fx["converged"]:
        o2 = Oracle(_solver_text(fx, o, cv["tolerance"]))
        x, it, res, st = o2.solve(np.zeros(o2.nvar))
        assert st == 0 and it == int(cv["iterations"])
        assert rel_err(o2.results()["free_energy"], cv["free_energy"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(SURFACE))
def test_gpu_surface_matches_reference(case):
    from namics_amd import ScfSystem
    from tests.test_gpu_parity import TOL_ORC, TOL_REF, check_results
    fx = _fixture(case)
    o = Oracle(fx["text"])
    print("surface case", case, flush=True)
    s = ScfSystem(fx["text"])
    assert s.layout()["mode"] == "fused"
    for ev in fx["evaluation"]:
        x = np.array(ev["x"])
        g, res = s.residual(x)
        check_results(s, s.results(), ev, TOL_REF)
        assert rel_err(res, ev["residual"]) < TOL_REF
        go, _ = o.residual(x)
        assert scaled_err(g, go) < TOL_ORC
    s.close()
    for cv in fx["converged"]:
        s = ScfSystem(_solver_text(fx, o, cv["tolerance"]))
        x, it, res, st = s.solve(np.zeros(s.nvar))
        print("surface case", case, "iterations", it, "reference", cv["iterations"], flush=True)
        assert st == 0 and res <= cv["tolerance"]
        r = s.results()
        assert rel_err(r["free_energy"], cv["free_energy"]) < max(1e-10, 100 * cv["tolerance"])
        for sn, p in cv["seg_phi"].items():
            assert scaled_err(s.interior(r["seg_phi"][sn]), p) < max(1e-10, 100 * cv["tolerance"])
        s.close()
