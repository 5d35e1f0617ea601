"""Tagged segments against the reference's own outputs.

A tagged segment has no field: its weight is its mask (Segment::
PrepareForCalculations, src/segment.cpp:859-879) and its density is the
mask (System::ComputePhis, src/system.cpp:2596-2601).  The molecule that
holds it has n = the number of tagged sites (src/system.cpp:606-611), is
normalised like a pinned one (src/system.cpp:2406-2416) and counts N - 1
segments in F, Omega and mu, with its other fractions scaled by N / (N - 1)
(src/system.cpp:2861-3057, 3262-3340).  The oracle restates that and
matches the reference fixtures (tests/golden/make_tagged_fixtures.py): F,
Omega, every molecule's GN, n, theta and mu, every density profile, and the
DIIS iteration counts; the library matches the same fixtures on the GPU.
"""
import json
import os

import numpy as np
import pytest

from tests.golden.tagged_cases import TAGGED
from tests.helpers import rel_err, scaled_err
from tests.oracle_lib import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(name):
    with open(os.path.join(HERE, "golden", "tagged", name + ".json")) as f:
        return json.load(f)


def _solver_text(fx, o, tol):
    return fx["text"] + "\nnewton : %s : method : DIIS\nnewton : %s : tolerance : %r\n" \
        "newton : %s : iterationlimit : 5000\n" % (o.newton_name, o.newton_name, tol, o.newton_name)


def _check(o, r, ref, tol):
    assert rel_err(r["free_energy"], ref["free_energy"]) < tol
    assert rel_err(r["grand_potential"], ref["grand_potential"]) < tol
    for m, d in ref["mol"].items():
        for k in ("GN", "n", "theta", "phibulk", "Mu"):
            assert abs(r["mol"][m][k] - d[k]) <= tol * max(1e-300, abs(d[k])), (m, k, r["mol"][m][k], d[k])
    for s, p in ref["seg_phi"].items():
        assert scaled_err(o.interior(r["seg_phi"][s]), p) < tol, s


@pytest.mark.parametrize("case", list(TAGGED))
def test_oracle_tagged_matches_reference(case):
    fx = _fixture(case)
    o = Oracle(fx["text"])
    assert o.nvar == len(fx["evaluation"][0]["x"])
    for ev in fx["evaluation"]:
        g, res = o.residual(np.array(ev["x"]))
        assert rel_err(res, ev["residual"]) < 1e-12
        _check(o, o.results(), ev, 1e-12)
    for cv in fx["converged"]:
        o2 = Oracle(_solver_text(fx, o, cv["tolerance"]))
        x, it, res, st = o2.solve(np.zeros(o2.nvar))
        assert st == 0 and it == int(cv["iterations"])
        _check(o2, o2.results(), cv, 1e-10)


def test_tagged_density_is_the_mask():
    """The tagged segment's density is 1 on its site and 0 elsewhere, and the
    molecule holds one chain per tagged site: theta = N."""
    fx = _fixture("tag1d_middle")
    o = Oracle(fx["text"])
    o.residual(np.array(fx["evaluation"][0]["x"]))
    r = o.results()
    phx = o.interior(r["seg_phi"]["X"])
    assert phx[7] == 1.0 and np.count_nonzero(phx) == 1
    assert r["mol"]["tag"]["n"] == 1.0 and abs(r["mol"]["tag"]["theta"] - 25.0) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(TAGGED))
def test_gpu_tagged_matches_reference(case):
    from namics_amd import ScfSystem
    from tests.test_gpu_parity import TOL_ORC, TOL_REF, check_results
    fx = _fixture(case)
    o = Oracle(fx["text"])
    print("tagged case", case, flush=True)
    s = ScfSystem(fx["text"])
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
        print("tagged case", case, "iterations", it, "reference", cv["iterations"], flush=True)
        assert st == 0 and res <= cv["tolerance"]
        _check_converged(s, s.results(), cv, max(1e-10, 100 * cv["tolerance"]))
        s.close()


def _check_converged(sysobj, r, ref, tol):
    """Converged state against the reference.  Where a tagged molecule is
    the only one that reaches some sites (the brush), the field there has a
    gauge: shifting it changes GN by a constant factor and leaves every
    density and F unchanged, so GN, mu and Omega depend on the path the
    iteration took (the oracle started from 1e-12 noise instead of zeros
    ends at a GN many decades away).  Those are compared in their
    gauge-free forms, mu + ln GN and Omega + n mu of the tagged molecules."""
    assert rel_err(r["free_energy"], ref["free_energy"]) < tol
    gp, gp_ref = r["grand_potential"], ref["grand_potential"]
    for m, d in ref["mol"].items():
        for k in ("n", "theta", "phibulk"):
            assert rel_err(r["mol"][m][k], d[k]) < tol, (m, k)
        mu, mu_ref = r["mol"][m]["Mu"], d["Mu"]
        if m in TAGGED_MOLS:
            mu, mu_ref = mu + np.log(r["mol"][m]["GN"]), mu_ref + np.log(d["GN"])
            gp, gp_ref = gp + d["n"] * r["mol"][m]["Mu"], gp_ref + d["n"] * d["Mu"]
        else:
            assert rel_err(r["mol"][m]["GN"], d["GN"]) < tol, (m, "GN")
        assert abs(mu - mu_ref) < tol * max(1.0, abs(mu_ref)), (m, "Mu")
    assert abs(gp - gp_ref) < tol * (abs(ref["free_energy"]) + abs(gp_ref))
    for sn, p in ref["seg_phi"].items():
        assert scaled_err(sysobj.interior(r["seg_phi"][sn]), p) < tol, sn


TAGGED_MOLS = ("brush", "tag")


def test_tagged_gauge():
    """The tagged brush's GN is fixed only up to the gauge: the oracle from
    two starts ends at GN many decades apart with the same F, densities,
    mu + ln GN and Omega + n mu."""
    fx = _fixture("tag1d_brush")
    o = Oracle(fx["text"])
    cv = fx["converged"][0]
    o2 = Oracle(_solver_text(fx, o, cv["tolerance"]))
    x, it, res, st = o2.solve(np.random.default_rng(1).uniform(-1e-8, 1e-8, o2.nvar))
    assert st == 0
    r = o2.results()
    assert abs(np.log10(r["mol"]["brush"]["GN"]) - np.log10(cv["mol"]["brush"]["GN"])) > 10
    _check_converged(o2, r, cv, 1e-6)
